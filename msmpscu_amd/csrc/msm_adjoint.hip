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
#include <algorithm>

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

// dL/dF in the sorted order as one 32-byte record per atom (x, y, z, 0), so that the pair pass
// gathers a neighbour's three components in one load; zero for inactive atoms.
__global__ void k_adj_gather(const double* __restrict__ dldf, const int* __restrict__ gid,
                             const int* __restrict__ statu, int n, double4* __restrict__ g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int q = gid[p];
    const bool act = (statu[p] & STATU_ACTIVE) == STATU_ACTIVE;
    const size_t m = static_cast<size_t>(n);
    g[p] = act ? make_double4(dldf[q], dldf[m + q], dldf[2 * m + q], 0.0)
               : make_double4(0.0, 0.0, 0.0, 0.0);
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
        fp = fs_fprime(rho);
        dfp = 0.25 / (rho * sqrt(rho));
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
    const double4* __restrict__ g4 = reinterpret_cast<const double4*>(d.g);
    for (int c = c0; c < c1; ++c) {
        const int gt = c * blockDim.x + threadIdx.x;
        const int k = gt >> 1, sub = gt & 1;
        const bool have = k < a.n;
        const bool active = have && (a.statu[k] & STATU_ACTIVE) == STATU_ACTIVE;
        double lam = 0.0;
        if (active) {
            const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, k);
            const double4 pk = a.pos[k];
            const double4 gk = g4[k];
            const double gkx = gk.x, gky = gk.y, gkz = gk.z;
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
                const double4 gj = g4[j];
                const double aj = gj.x * sx + gj.y * sy + gj.z * sz;
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
                    AdjDev d, double* out_pairs, double* out_dens, hipStream_t s,
                    bool rho_ready) {
    if (a.n <= 0) return;
    const bool eam = t.pot_type != 0, mt = t.ngroup > 1;
    const unsigned g1 = static_cast<unsigned>((a.n + 255) / 256);
    hipLaunchKernelGGL(k_adj_gather, dim3(g1), dim3(256), 0, s, dldf, gid, a.statu, a.n,
                       reinterpret_cast<double4*>(d.g));
#define MSM_AR(MB, E, M) hipLaunchKernelGGL((k_adj_rho<MB, E, M>), dim3(g1), dim3(256), 0, s, a, \
                                            l, t, box0, boxes, d)
#define MSM_ARM(MB, E) do { if (mt) MSM_AR(MB, E, 1); else MSM_AR(MB, E, 0); } while (0)
    if (!rho_ready) {
        if (multibox) { if (eam) MSM_ARM(1, 1); else MSM_ARM(1, 0); }
        else          { if (eam) MSM_ARM(0, 1); else MSM_ARM(0, 0); }
    }
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

// --------------------------------------------------------------------------------------------
// Device-resident force matching: the tables of the cubic B-spline coefficient model, the loss
// with the inputs of the adjoint, and the coefficient gradient (the transpose of the table map).
// The arithmetic is operation for operation that of the host model (adjoint.py SplineModel), with
// contraction off, so the tables and gradients equal the host's bit for bit.
// --------------------------------------------------------------------------------------------

namespace {

// Value and slope at knot i (1..n) of a unit-spacing cubic B-spline with coefficients c[0..n-1]
// (zero outside): ((c[i-2] + 4 c[i-1]) + c[i]) / 6 and (c[i] - c[i-2]) / 2; zero for knots
// outside 1..n (the tables' padding).
__device__ __forceinline__ double bs_val(const double* c, int n, int i) {
#pragma clang fp contract(off)
    if (i < 1 || i > n) return 0.0;
    const double a = i >= 2 ? c[i - 2] : 0.0, d = i < n ? c[i] : 0.0;
    return ((a + 4.0 * c[i - 1]) + d) / 6.0;
}

__device__ __forceinline__ double bs_slope(const double* c, int n, int i) {
#pragma clang fp contract(off)
    if (i < 1 || i > n) return 0.0;
    const double a = i >= 2 ? c[i - 2] : 0.0, d = i < n ? c[i] : 0.0;
    return 0.5 * (d - a);
}

// The spline tables at knot i of one kind (zero outside 1..ntab / 1..nembd):
//   POTR = v r, FPOTR = -i s, POTB = v, FPOTB = ((-s) CSI^2) / (2i), r = (i/CSI)^2,
//   FEMBD = v, DFEMBD = s / RHOD.
struct SplineTab {
    const double *cp, *cq, *ce;
    int ntab, nembd;
    double csi, csi2, rhod;
    __device__ double potr(int i) const {
#pragma clang fp contract(off)
        if (i < 1 || i > ntab) return 0.0;
        const double x = static_cast<double>(i) / csi;
        return bs_val(cp, ntab, i) * (x * x);
    }
    __device__ double fpotr(int i) const {
#pragma clang fp contract(off)
        if (i < 1 || i > ntab) return 0.0;
        return -static_cast<double>(i) * bs_slope(cp, ntab, i);
    }
    __device__ double potb(int i) const { return bs_val(cq, ntab, i); }
    __device__ double fpotb(int i) const {
#pragma clang fp contract(off)
        if (i < 1 || i > ntab) return 0.0;
        return (-bs_slope(cq, ntab, i) * csi2) / (2.0 * static_cast<double>(i));
    }
    __device__ double fembd(int i) const { return bs_val(ce, nembd, i); }
    __device__ double dfembd(int i) const {
#pragma clang fp contract(off)
        if (i < 1 || i > nembd) return 0.0;
        return bs_slope(ce, nembd, i) / rhod;
    }
};

// One thread per padded knot a: the device table entries msm_set_force_table builds on the host,
// (T(a), T(a+1) - T(a)) pairs, from the spline tables.
__global__ void k_spline_tables(SplineTab st, double2* __restrict__ b2, double4* __restrict__ f4,
                                double4* __restrict__ e4, double4* __restrict__ m4) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < st.ntab + 2) {
        const double pb = st.potb(a), pb1 = st.potb(a + 1);
        const double fr = st.fpotr(a), fr1 = st.fpotr(a + 1);
        const double fb = st.fpotb(a), fb1 = st.fpotb(a + 1);
        const double pr = st.potr(a), pr1 = st.potr(a + 1);
        b2[a] = make_double2(pb, pb1 - pb);
        f4[a] = make_double4(fr, fr1 - fr, fb, fb1 - fb);
        e4[a] = make_double4(pr, pr1 - pr, pb, pb1 - pb);
    }
    if (m4 && a < st.nembd + 2) {
        const double f0 = st.fembd(a), f1 = st.fembd(a + 1);
        const double d0 = st.dfembd(a), d1 = st.dfembd(a + 1);
        m4[a] = make_double4(f0, f1 - f0, d0, d1 - d0);
    }
}

// Per sorted atom p (caller index gid[p]): dF = F_p - F*, dL/dF = (2 dF) / N into the caller's
// order and FP(N,3) layout (the adjoint's input), and |dF|^2 summed per workgroup in a fixed
// order (grid-stride partials, then a fixed-shape tree).
__global__ __launch_bounds__(kAdjBlock) void k_fm_forces(const double* __restrict__ fx,
                                                         const double* __restrict__ fy,
                                                         const double* __restrict__ fz,
                                                         const int* __restrict__ gid, int n,
                                                         const double* __restrict__ ft,
                                                         double* __restrict__ dldf,
                                                         double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double lds[kAdjBlock];
    double acc = 0.0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const size_t c = static_cast<size_t>(gid[p]);
        const double dx = fx[p] - ft[c], dy = fy[p] - ft[c + n], dz = fz[p] - ft[c + 2 * static_cast<size_t>(n)];
        const double nd = static_cast<double>(n);
        dldf[c] = (2.0 * dx) / nd;
        dldf[c + n] = (2.0 * dy) / nd;
        dldf[c + 2 * static_cast<size_t>(n)] = (2.0 * dz) / nd;
        acc = acc + ((dx * dx + dy * dy) + dz * dz);
    }
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kAdjBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = lds[0];
}

// One workgroup: L = (sum of the force partials) / N + w_e sum_b ((E_b - E*_b)/N_b)^2 / NBOX and
// dL/dE_b = 2 w_e (E_b - E*_b) / (N_b^2 NBOX), each sum in a fixed order.
__global__ __launch_bounds__(kAdjBlock) void k_fm_final(const double* __restrict__ part, int np,
                                                        const double* __restrict__ ebox,
                                                        const double* __restrict__ et,
                                                        const BoxDev* __restrict__ boxes,
                                                        int nbox, int n, double w_e,
                                                        double* __restrict__ gbox,
                                                        double* __restrict__ loss) {
#pragma clang fp contract(off)
    __shared__ double lf[kAdjBlock], le[kAdjBlock];
    double sf = 0.0, se = 0.0;
    for (int k = threadIdx.x; k < np; k += kAdjBlock) sf = sf + part[k];
    for (int b = threadIdx.x; b < nbox; b += kAdjBlock) {
        const double nb = static_cast<double>(boxes[b].natoms);
        const double d = ebox[b] - et[b];
        const double de = d / nb;
        gbox[b] = 2.0 * w_e * d / (nb * nb * static_cast<double>(nbox));
        se = se + de * de;
    }
    lf[threadIdx.x] = sf;
    le[threadIdx.x] = se;
    __syncthreads();
    for (int w = kAdjBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            lf[threadIdx.x] += lf[threadIdx.x + w];
            le[threadIdx.x] += le[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *loss = lf[0] / static_cast<double>(n) + w_e * le[0] / static_cast<double>(nbox);
}

// Knot gradient y(k), k = 1..n, from a bin row (zero outside 1..n).
__device__ __forceinline__ double bin_at(const double* row, int n, int k) {
    return (k < 1 || k > n) ? 0.0 : row[k];
}

// dL/dc of coefficient j of each row, the transpose of the table map:
//   row 0: V^T(POTR' r) + S^T(-i FPOTR'), row 1: V^T(POTB') + S^T(((-FPOTB') CSI^2) / (2i)),
//   row 2: V^T(FEMBD') + S^T(DFEMBD' / RHOD),
// with V^T(y)_j = ((y(j) + 4 y(j+1)) + y(j+2)) / 6 and S^T(y)_j = (y(j) - y(j+2)) / 2 over the
// knots k = 1..n (row 2 zero for FS tables).
__global__ void k_spline_grad(SplineTab st, const double* __restrict__ pairs,
                              const double* __restrict__ dens, int nt, int ne, int eam, int ld,
                              double* __restrict__ g) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ld) return;
    const double* g_fpotr = pairs;
    const double* g_fpotb = pairs + nt;
    const double* g_potr = pairs + 2 * static_cast<size_t>(nt);
    const double* g_potb = dens;
    const int n = st.ntab;
    auto yr = [&](int k) {
        if (k < 1 || k > n) return 0.0;
        const double x = static_cast<double>(k) / st.csi;
        return bin_at(g_potr, n, k) * (x * x);
    };
    auto yf = [&](int k) { return (k < 1 || k > n) ? 0.0 : -static_cast<double>(k) * g_fpotr[k]; };
    auto yb = [&](int k) { return bin_at(g_potb, n, k); };
    auto yfb = [&](int k) {
        return (k < 1 || k > n) ? 0.0 : (-g_fpotb[k] * st.csi2) / (2.0 * static_cast<double>(k));
    };
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    if (j < n) {
        r0 = ((yr(j) + 4.0 * yr(j + 1)) + yr(j + 2)) / 6.0 + 0.5 * (yf(j) - yf(j + 2));
        r1 = ((yb(j) + 4.0 * yb(j + 1)) + yb(j + 2)) / 6.0 + 0.5 * (yfb(j) - yfb(j + 2));
    }
    if (eam && j < st.nembd) {
        const int m = st.nembd;
        const double* g_fembd = dens + nt;
        const double* g_dfembd = dens + nt + ne;
        auto ye = [&](int k) { return bin_at(g_fembd, m, k); };
        auto yd = [&](int k) { return (k < 1 || k > m) ? 0.0 : g_dfembd[k] / st.rhod; };
        r2 = ((ye(j) + 4.0 * ye(j + 1)) + ye(j + 2)) / 6.0 + 0.5 * (yd(j) - yd(j + 2));
    }
    g[j] = r0;
    g[ld + j] = r1;
    g[2 * static_cast<size_t>(ld) + j] = r2;
}

}  // namespace

void launch_spline_tables(const double* coef, int ld, const TabDev& t, double2* b2, double4* f4,
                          double4* e4, double4* m4, hipStream_t s) {
    SplineTab st{coef, coef + ld, coef + 2 * static_cast<size_t>(ld), t.ntab2 - 2,
                 t.pot_type != 0 ? t.nembd : 0, t.csi, t.csi * t.csi, t.rhod};
    const int m = std::max(t.ntab2, t.pot_type != 0 ? t.nembd2 : 0);
    hipLaunchKernelGGL(k_spline_tables, dim3((m + 255) / 256), dim3(256), 0, s, st, b2, f4, e4,
                       t.pot_type != 0 ? m4 : nullptr);
}

int fm_parts(int n) { return std::max(1, std::min(512, (n + kAdjBlock - 1) / kAdjBlock)); }

void launch_fm_loss(const double* fx, const double* fy, const double* fz, const int* gid, int n,
                    const double* ft, const double* ebox, const double* et, const BoxDev* boxes,
                    int nbox, double w_e, double* dldf, double* gbox, double* part, double* loss,
                    hipStream_t s) {
    const int np = fm_parts(n);
    hipLaunchKernelGGL(k_fm_forces, dim3(np), dim3(kAdjBlock), 0, s, fx, fy, fz, gid, n, ft,
                       dldf, part);
    hipLaunchKernelGGL(k_fm_final, dim3(1), dim3(kAdjBlock), 0, s, part, np, ebox, et, boxes, nbox,
                       n, w_e, gbox, loss);
}

void launch_spline_grad(const double* pairs, const double* dens, const TabDev& t, int ld,
                        double* g, hipStream_t s) {
    const bool eam = t.pot_type != 0;
    SplineTab st{nullptr, nullptr, nullptr, t.ntab2 - 2, eam ? t.nembd : 0, t.csi, t.csi * t.csi,
                 t.rhod};
    hipLaunchKernelGGL(k_spline_grad, dim3((ld + 255) / 256), dim3(256), 0, s, st, pairs, dens,
                       t.ntab2, eam ? t.nembd2 : 0, eam ? 1 : 0, ld, g);
}

}  // namespace msm
