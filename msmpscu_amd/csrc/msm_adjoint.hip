// msm_adjoint.hip -- reverse-mode derivative of a loss of the forces and box energies with
// respect to the knots of the interpolation tables (every table kind).
//
// Forward model (the density/force/energy kernels; MD_FS_ForceTable_GPU.F90 PRECALFOR_KERNEL,
// CALFORCE_KERNEL, CALEPOT_KERNEL and their EAM twins), for active atoms k and in-cutoff pairs:
//   rho_k   = sum_j B(r_kj),                     B = lerp(POTB)
//   F'_k    = F'(rho_k)   (FS: -1/(2 sqrt rho); EAM: lerp(DFEMBD) at rho/RHOD + 1; rho itself
//                          when rho <= 0, as the reference leaves it)
//   F_k     = sum_j s_kj [R'(r)/r^2 + Bp(r) (F'_k + F'_j)/r],   R' = lerp(FPOTR), Bp = lerp(FPOTB)
//   E_k     = sum_j R(r)/r + F(rho_k)            R = lerp(POTR); FS: F = -sqrt(rho); EAM: lerp(FEMBD)
//   E_b     = sum_{k in box b} E_k
// With several atom types a pair (k, j) reads the kind KT = KPAIR(t_k, t_j) for B, R', R and the
// Bp that multiplies F'_k, the kind KT1 = KPAIR(t_j, t_k) for the Bp that multiplies F'_j, and
// the embedding tables of kind KEMBD(t_k).  Every table enters through lerp weights (1-dk, dk) on
// knots (KK, KK+1), so with G_k = dL/dF_k (zero for inactive atoms) and g_b = dL/dE_b the knot
// gradients are sums of weighted per-pair terms:
//   dL/dFPOTR[KT]  += (G_k . s_kj) / r^2
//   dL/dFPOTB[KT]  += (G_k . s_kj) F'_k / r,   dL/dFPOTB[KT1] += (G_k . s_kj) F'_j / r
//   dL/dPOTR[KT]   += g_b(k) / r
//   lambda_k    = dL/dF'_k = sum_j (G_k - G_j) . s_kj Bp[KT] / r   (F'_k meets Bp[KT] in both
//                 F_k and F_j)
//   mu_k        = dL/drho_k = lambda_k dF'/drho + g_b(k) dF/drho
//   dL/dPOTB[KT] += mu_k,  dL/dDFEMBD[KEMBD(t_k)] += lambda_k (rho_k > 0),
//   dL/dFEMBD[KEMBD(t_k)] += g_b(k)
// Three passes (rho; pair terms + lambda; density terms + mu) over the list, each adding its
// terms into per-workgroup bins in LDS (ds_add_f64), then an ordered sum of the workgroups'
// bins.  The bins of one workgroup see their terms in an order that can vary from run to run,
// so the result is reproducible to rounding, not bit for bit.
#include "msm_device.h"
#include "msm_internal.h"

namespace msm {

namespace {

constexpr int kAdjBlock = 1024;

__device__ __forceinline__ void bin_add(double* bins, int kk, double dk, double w) {
    atomicAdd(&bins[kk], w * (1.0 - dk));
    atomicAdd(&bins[kk + 1], w * dk);
}

// workgroup range of chunks (contiguous per workgroup)
__device__ __forceinline__ void adj_range(int nchunks, int& c0, int& c1) {
    const int nblk = gridDim.x, L = blockIdx.x;
    c0 = static_cast<int>(static_cast<long long>(L) * nchunks / nblk);
    c1 = static_cast<int>(static_cast<long long>(L + 1) * nchunks / nblk);
}

// bins of this workgroup: LDS, or its slice of the partial buffer when the tables are too big
template <bool LDSB>
__device__ __forceinline__ double* adj_bins(double* lds, const AdjDev& d, int nbin) {
    double* p = LDSB ? lds : d.part + static_cast<size_t>(blockIdx.x) * d.part_stride;
    for (int k = threadIdx.x; k < nbin; k += blockDim.x) p[k] = 0.0;
    __syncthreads();
    return p;
}

template <bool LDSB>
__device__ __forceinline__ void adj_flush(const double* bins, const AdjDev& d, int nbin) {
    __syncthreads();
    if constexpr (LDSB) {
        double* out = d.part + static_cast<size_t>(blockIdx.x) * d.part_stride;
        for (int k = threadIdx.x; k < nbin; k += blockDim.x) out[k] = bins[k];
    }
}

__device__ __forceinline__ double pair_sum2(double v) { return v + __shfl_xor(v, 1, 64); }

}  // namespace

// dL/dF in the sorted order, zero for inactive atoms.
__global__ void k_adj_gather(const double* __restrict__ dldf, const int* __restrict__ gid,
                             const int* __restrict__ statu, int n, double* __restrict__ g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int q = gid[p];
    const bool act = (statu[p] & STATU_ACTIVE) == STATU_ACTIVE;
    for (int c = 0; c < 3; ++c) g[c * static_cast<size_t>(n) + p] = act ? dldf[c * static_cast<size_t>(n) + q] : 0.0;
}

template <int EAM>
__device__ __forceinline__ void fprime(const TabDev& t, int ke, double rho, double& fp,
                                       double& dfp);

// rho_k and F'_k (same arithmetic as the density pass; F' goes to pos.w as there).
template <int MB, int EAM, int MT>
__global__ void k_adj_rho(AtomsDev a, ListDev l, TabDev t, BoxDev box0,
                          const BoxDev* __restrict__ boxes, AdjDev d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double den0 = 0.0;
    const bool active = (a.statu[i] & STATU_ACTIVE) == STATU_ACTIVE;
    if (active) {
        const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, i);
        const double4 pi = a.pos[i];
        const int kv = l.kvois[i];
        const int ti = MT ? a.ityp[i] : 0;
        for (int iw = 0; iw < kv; ++iw) {
            const int j = l.indi[list_base(i, l.mx) + static_cast<size_t>(iw) * l.stride];
            const double4 pj = a.pos[j];
            const double sx = min_image(pi.x - pj.x, b.len[0], b.ilen[0]);
            const double sy = min_image(pi.y - pj.y, b.len[1], b.ilen[1]);
            const double sz = min_image(pi.z - pj.z, b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            if (r2 <= t.ru2max) {
                const int kind = t.kpair[MT ? ti * t.ngroup + a.ityp[j] : 0];
                double inv_r;
                const double sk = table_coord(r2, t.csi, inv_r);
                const int kk = min(static_cast<int>(sk), t.ntab2 - 2);
                const double2 e = t.b2[kind * t.ntab2 + kk];
                den0 = den0 + (e.x + (sk - kk) * e.y);
            }
        }
    }
    d.rho[i] = den0;
    double fp = 0.0, dfp;
    if (active) fprime<EAM>(t, t.kembd[MT ? a.ityp[i] : 0], den0, fp, dfp);
    a.pos[i].w = fp;
}

// F'(rho) and its derivative (the density pass's finishing step).
template <int EAM>
__device__ __forceinline__ void fprime(const TabDev& t, int ke, double rho, double& fp,
                                       double& dfp) {
    if (!(rho > 0.0)) {
        fp = rho;  // the reference leaves DEN = rho when rho <= 0
        dfp = 1.0;
        return;
    }
    if constexpr (!EAM) {
        const double sr = sqrt(rho);
        fp = -0.5 / sr;
        dfp = 0.25 / (rho * sr);
    } else {
        const double sk = rho / t.rhod + 1.0;
        int kk = static_cast<int>(sk + 0.000001);
        if (kk > t.nembd) kk = t.nembd;
        const double4 m = t.m4[ke * t.nembd2 + kk];
        fp = m.z + (sk - kk) * m.w;
        dfp = m.w / t.rhod;
    }
}

// Pair pass: FPOTR / FPOTB / POTR knot terms and lambda_k.  Bins: [FPOTR | FPOTB | POTR], each
// nkind x ntab2, kind-major.
template <int MB, int EAM, bool LDSB, int MT>
__global__ __launch_bounds__(kAdjBlock) void k_adj_pairs(AtomsDev a, ListDev l, TabDev t,
                                                         BoxDev box0,
                                                         const BoxDev* __restrict__ boxes,
                                                         AdjDev d, int nchunks) {
    extern __shared__ double s_bins[];
    const int nt = t.ntab2, nkt = (MT ? t.nkind : 1) * nt;
    double* bins = adj_bins<LDSB>(s_bins, d, 3 * nkt);
    double* b_fr = bins;
    double* b_fb = bins + nkt;
    double* b_pr = bins + 2 * nkt;
    int c0, c1;
    adj_range(nchunks, c0, c1);
    const size_t n = static_cast<size_t>(a.n);
    for (int c = c0; c < c1; ++c) {
        const int gt = c * blockDim.x + threadIdx.x;
        const int k = gt >> 1, sub = gt & 1;
        const bool have = k < a.n;
        const bool active = have && (a.statu[k] & STATU_ACTIVE) == STATU_ACTIVE;
        double lam = 0.0;
        if (active) {
            const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, k);
            const double4 pk = a.pos[k];
            const double gkx = d.g[k], gky = d.g[n + k], gkz = d.g[2 * n + k];
            const double gb = d.gbox[MB ? a.boxid[k] : 0];
            const int kv = l.kvois[k];
            const int tk = MT ? a.ityp[k] : 0;
            for (int iw = sub; iw < kv; iw += 2) {
                const int j = l.indi[list_base(k, l.mx) + static_cast<size_t>(iw) * l.stride];
                const double4 pj = a.pos[j];
                const double sx = min_image(pk.x - pj.x, b.len[0], b.ilen[0]);
                const double sy = min_image(pk.y - pj.y, b.len[1], b.ilen[1]);
                const double sz = min_image(pk.z - pj.z, b.len[2], b.ilen[2]);
                const double r2 = sx * sx + sy * sy + sz * sz;
                if (!(r2 <= t.ru2max)) continue;
                int kind = t.kpair[0], kind1 = t.kpair[0];
                if constexpr (MT) {
                    const int tj = a.ityp[j];
                    kind = t.kpair[tk * t.ngroup + tj];
                    kind1 = t.kpair[tj * t.ngroup + tk];
                }
                double inv_r;
                const double sk = table_coord(r2, t.csi, inv_r);
                const int kk = min(static_cast<int>(sk), t.ntab2 - 2);
                const double dk = sk - kk;
                const double4 f = t.f4[kind * nt + kk];
                const double bp = f.z + dk * f.w;
                const double ak = gkx * sx + gky * sy + gkz * sz;
                const double aj = d.g[j] * sx + d.g[n + j] * sy + d.g[2 * n + j] * sz;
                const int row = MT ? kind * nt : 0;
                bin_add(b_fr + row, kk, dk, ak * inv_r * inv_r);
                if constexpr (MT) {
                    bin_add(b_fb + row, kk, dk, ak * pk.w * inv_r);
                    bin_add(b_fb + kind1 * nt, kk, dk, ak * pj.w * inv_r);
                } else {
                    bin_add(b_fb, kk, dk, ak * (pk.w + pj.w) * inv_r);
                }
                bin_add(b_pr + row, kk, dk, gb * inv_r);
                lam = lam + (ak - aj) * bp * inv_r;
            }
        }
        lam = pair_sum2(lam);
        if (have && sub == 0) d.lam[k] = active ? lam : 0.0;
    }
    adj_flush<LDSB>(bins, d, 3 * nkt);
}

// Density pass: mu_k, POTB knot terms and the embedding knot terms.  Bins: [POTB | FEMBD |
// DFEMBD] (nkind x ntab2, then nkind1 x nembd2 twice; the embedding parts only for EAM).
template <int MB, int EAM, bool LDSB, int MT>
__global__ __launch_bounds__(kAdjBlock) void k_adj_density(AtomsDev a, ListDev l, TabDev t,
                                                           BoxDev box0,
                                                           const BoxDev* __restrict__ boxes,
                                                           AdjDev d, int nchunks) {
    extern __shared__ double s_bins[];
    const int nt = t.ntab2, ne = EAM ? t.nembd2 : 0;
    const int nkt = (MT ? t.nkind : 1) * nt, nke = (MT ? t.nkind1 : 1) * ne;
    double* bins = adj_bins<LDSB>(s_bins, d, nkt + 2 * nke);
    double* b_pb = bins;
    double* b_fe = bins + nkt;
    double* b_df = bins + nkt + nke;
    int c0, c1;
    adj_range(nchunks, c0, c1);
    for (int c = c0; c < c1; ++c) {
        const int gt = c * blockDim.x + threadIdx.x;
        const int k = gt >> 1, sub = gt & 1;
        const bool have = k < a.n;
        const bool active = have && (a.statu[k] & STATU_ACTIVE) == STATU_ACTIVE;
        if (!active) continue;
        const double rho = d.rho[k];
        const double lam = d.lam[k];
        const double gb = d.gbox[MB ? a.boxid[k] : 0];
        const int tk = MT ? a.ityp[k] : 0;
        const int ke = t.kembd[tk];
        double fp, dfp;
        fprime<EAM>(t, ke, rho, fp, dfp);
        double de;  // dF/drho of the energy term
        if constexpr (!EAM) {
            de = rho > 0.0 ? fs_fprime(rho) : 0.0;
        } else {
            // the energy kernel's lookup: KK = int(rho/RHOD + 1 + 1e-6) clamped to [0, NEMBD]
            const double sk = rho / t.rhod + 1.0;
            int kk = static_cast<int>(sk + 0.000001);
            if (kk < 0) kk = 0;
            if (kk > t.nembd) kk = t.nembd;
            const double4 m = t.m4[ke * t.nembd2 + kk];
            de = m.y / t.rhod;
            if (sub == 0) {
                const int erow = MT ? ke * ne : 0;
                bin_add(b_fe + erow, kk, sk - kk, gb);
                if (rho > 0.0) bin_add(b_df + erow, kk, sk - kk, lam);
            }
        }
        const double mu = lam * dfp + gb * de;
        const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, k);
        const double4 pk = a.pos[k];
        const int kv = l.kvois[k];
        for (int iw = sub; iw < kv; iw += 2) {
            const int j = l.indi[list_base(k, l.mx) + static_cast<size_t>(iw) * l.stride];
            const double4 pj = a.pos[j];
            const double sx = min_image(pk.x - pj.x, b.len[0], b.ilen[0]);
            const double sy = min_image(pk.y - pj.y, b.len[1], b.ilen[1]);
            const double sz = min_image(pk.z - pj.z, b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            if (!(r2 <= t.ru2max)) continue;
            const int row = MT ? t.kpair[tk * t.ngroup + a.ityp[j]] * nt : 0;
            double inv_r;
            const double sk = table_coord(r2, t.csi, inv_r);
            const int kk = min(static_cast<int>(sk), t.ntab2 - 2);
            bin_add(b_pb + row, kk, sk - kk, mu);
        }
    }
    adj_flush<LDSB>(bins, d, nkt + 2 * nke);
}

// out[k] = sum over workgroups w of part[w * stride + k], k < nbin, in a fixed order: a
// workgroup takes kRedBins consecutive bins; its kRedSplit thread rows each sum the partials of
// every kRedSplit-th workgroup (eight loads in flight per thread), and the rows' sums meet in a
// fixed-shape tree in LDS.
constexpr int kRedBins = 64, kRedSplit = 4;

__global__ __launch_bounds__(kRedBins * kRedSplit) void k_adj_reduce(
    const double* __restrict__ part, int nwg, size_t stride, int nbin, double* __restrict__ out) {
    __shared__ double acc[kRedSplit][kRedBins];
    const int kb = threadIdx.x % kRedBins, row = threadIdx.x / kRedBins;
    const int k = blockIdx.x * kRedBins + kb;
    double s = 0.0;
    if (k < nbin) {
        int w = row;
        for (; w + 7 * kRedSplit < nwg; w += 8 * kRedSplit) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = part[static_cast<size_t>(w + u * kRedSplit) * stride + k];
#pragma unroll
            for (int u = 0; u < 8; ++u) s = s + v[u];
        }
        for (; w < nwg; w += kRedSplit) s = s + part[static_cast<size_t>(w) * stride + k];
    }
    acc[row][kb] = s;
    __syncthreads();
    if (row == 0 && k < nbin) {
        double t = acc[0][kb];
        for (int r = 1; r < kRedSplit; ++r) t = t + acc[r][kb];
        out[k] = t;
    }
}

namespace {

int adj_lds_limit() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
        return 65536;
    return v;
}

int adj_cus() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v <= 0)
        return 256;
    return v;
}

}  // namespace

int adjoint_workgroups(int n) {
    const int nchunks = static_cast<int>((2LL * n + kAdjBlock - 1) / kAdjBlock);
    const int cus = adj_cus();
    return nchunks < cus ? (nchunks > 0 ? nchunks : 1) : cus;
}

// Bin counts of the two passes: [FPOTR | FPOTB | POTR] and [POTB | FEMBD | DFEMBD].
void adjoint_bins(const TabDev& t, int& nb1, int& nb2) {
    const bool mt = t.ngroup > 1, eam = t.pot_type != 0;
    const int nkt = (mt ? t.nkind : 1) * t.ntab2;
    const int nke = eam ? (mt ? t.nkind1 : 1) * t.nembd2 : 0;
    nb1 = 3 * nkt;
    nb2 = nkt + 2 * nke;
}

void launch_adjoint(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                    const BoxDev* boxes, int multibox, const double* dldf, const int* gid,
                    AdjDev d, double* out_pairs, double* out_dens, hipStream_t s) {
    if (a.n <= 0) return;
    const bool eam = t.pot_type != 0, mt = t.ngroup > 1;
    const unsigned g1 = static_cast<unsigned>((a.n + 255) / 256);
    hipLaunchKernelGGL(k_adj_gather, dim3(g1), dim3(256), 0, s, dldf, gid, a.statu, a.n, d.g);
#define MSM_AR(MB, E, M) hipLaunchKernelGGL((k_adj_rho<MB, E, M>), dim3(g1), dim3(256), 0, s, a, \
                                            l, t, box0, boxes, d)
#define MSM_ARM(MB, E) do { if (mt) MSM_AR(MB, E, 1); else MSM_AR(MB, E, 0); } while (0)
    if (multibox) { if (eam) MSM_ARM(1, 1); else MSM_ARM(1, 0); }
    else          { if (eam) MSM_ARM(0, 1); else MSM_ARM(0, 0); }
#undef MSM_ARM
#undef MSM_AR
    const int nchunks = static_cast<int>((2LL * a.n + kAdjBlock - 1) / kAdjBlock);
    const int nwg = adjoint_workgroups(a.n);
    const int lim = adj_lds_limit();
    int nb1, nb2;
    adjoint_bins(t, nb1, nb2);
    const bool l1 = static_cast<size_t>(nb1) * sizeof(double) <= static_cast<size_t>(lim);
    const bool l2 = static_cast<size_t>(nb2) * sizeof(double) <= static_cast<size_t>(lim);
    const size_t s1 = l1 ? nb1 * sizeof(double) : 0, s2 = l2 ? nb2 * sizeof(double) : 0;
#define MSM_AP(MB, E, LB, M)                                                               \
    hipLaunchKernelGGL((k_adj_pairs<MB, E, LB, M>), dim3(nwg), dim3(kAdjBlock), s1, s, a, l, \
                       t, box0, boxes, d, nchunks)
#define MSM_APL(MB, E)                                                              \
    do {                                                                            \
        if (mt) { if (l1) MSM_AP(MB, E, true, 1); else MSM_AP(MB, E, false, 1); }   \
        else    { if (l1) MSM_AP(MB, E, true, 0); else MSM_AP(MB, E, false, 0); }   \
    } while (0)
    if (multibox) { if (eam) MSM_APL(1, 1); else MSM_APL(1, 0); }
    else          { if (eam) MSM_APL(0, 1); else MSM_APL(0, 0); }
#undef MSM_APL
#undef MSM_AP
    const unsigned gr1 = static_cast<unsigned>((nb1 + kRedBins - 1) / kRedBins);
    hipLaunchKernelGGL(k_adj_reduce, dim3(gr1), dim3(kRedBins * kRedSplit), 0, s, d.part, nwg,
                       d.part_stride, nb1, out_pairs);
#define MSM_AD(MB, E, LB, M)                                                                 \
    hipLaunchKernelGGL((k_adj_density<MB, E, LB, M>), dim3(nwg), dim3(kAdjBlock), s2, s, a, l, \
                       t, box0, boxes, d, nchunks)
#define MSM_ADL(MB, E)                                                              \
    do {                                                                            \
        if (mt) { if (l2) MSM_AD(MB, E, true, 1); else MSM_AD(MB, E, false, 1); }   \
        else    { if (l2) MSM_AD(MB, E, true, 0); else MSM_AD(MB, E, false, 0); }   \
    } while (0)
    if (multibox) { if (eam) MSM_ADL(1, 1); else MSM_ADL(1, 0); }
    else          { if (eam) MSM_ADL(0, 1); else MSM_ADL(0, 0); }
#undef MSM_ADL
#undef MSM_AD
    const unsigned gr2 = static_cast<unsigned>((nb2 + kRedBins - 1) / kRedBins);
    hipLaunchKernelGGL(k_adj_reduce, dim3(gr2), dim3(kRedBins * kRedSplit), 0, s, d.part, nwg,
                       d.part_stride, nb2, out_dens);
}

}  // namespace msm
