// msm_ldstab.hip -- density and force passes with the interpolation table held in LDS.
//
// The per-pair table lookups of the global-list passes are scattered 16/32-byte loads from a
// 80-160 KB table per kind; together with the neighbour positions they overflow the 32 KB L1 and
// feed the vector-memory pipeline twice per pair.  Here each workgroup first copies a window of
// knots of the table kinds in use into LDS -- values only, {FPOTR, FPOTB} for the force pass
// (16 B/knot/kind) and POTB for the density pass (8 B/knot/kind) -- and the lerp reads knots KK
// and KK+1 from there:
//   T(KK) + DK * (T(KK+1) - T(KK))      (MD_FS_ForceTable_GPU.F90 PRECALFOR/CALFORCE kernels)
// which is the same arithmetic as the reference.  The window is the top of the table (large
// separations, where every neighbour lies); a rare pair below the window (r under about 1 A for
// the W tables) is flagged and its term added after its group of neighbours, from the knots of
// the global table.
//
// The grid is persistent: a chosen number of workgroups per CU (capped by what fits), each
// filling its LDS window once and then walking a contiguous range of atom chunks (contiguous per
// XCD), so the waves sharing a CU's L1 work on one compact region of the cell-sorted atoms.  Lanes per atom
// (TPA), list traversal and the finishing lane are those of msm_tpa.hip.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "msm_brick.h"
#include "msm_lanepair.h"

namespace msm {

namespace {

constexpr int kGroup = 8;
// Slots per software-pipelined group of the density and force passes, measured on the 1M-atom
// W box: density 4 / 6 / 8 took 86.7 / 86.2 / 87.6 us, force 3 / 4 took 105.7 / 107.2 us (8 for
// the force pass spills).
constexpr int kGroupD = 6;
constexpr int kGroupF = 3;

// TPA = 3 stands for two lanes per atom (lanes 2m, 2m + 1) sharing their position gathers.
template <int TPA>
__device__ __forceinline__ double lane_sum(double v) {
#pragma unroll
    for (int d = TPA / 2; d >= 1; d >>= 1) v = v + __shfl_xor(v, d, 64);
    return v;
}

// Slot stride of one lane: lanes per atom.
template <int TPA>
constexpr int slot_step() {
    return TPA == 3 ? 2 : TPA;
}

// Atom and lane-within-atom of this thread in chunk c: TPA consecutive lanes per atom; TPA = 3
// is the two-lane layout (lanes 2m, 2m + 1) with the pair's position gathers shared.
template <int TPA>
__device__ __forceinline__ void atom_of_lane(int c, int& i, int& sub) {
    constexpr int L = TPA == 3 ? 2 : TPA;
    const int gt = c * blockDim.x + threadIdx.x;
    i = gt / L;
    sub = gt % L;
}

// chunk range [c0, c1) of this workgroup
__device__ __forceinline__ void chunk_range(int nchunks, int& c0, int& c1) {
    const int nblk = gridDim.x;
    const int L = xcd_block(nblk);
    c0 = static_cast<int>(static_cast<long long>(L) * nchunks / nblk);
    c1 = static_cast<int>(static_cast<long long>(L + 1) * nchunks / nblk);
}

// Chunks c = first, first + step, ... < end of this workgroup.  sweep = 0: the contiguous range
// of chunk_range.  sweep = 1: XCD x (workgroup ids x, x+8, ...) owns the contiguous range
// [x*n/8, (x+1)*n/8) and its workgroups take its chunks round-robin (plain striding with fewer
// than 8 workgroups).  Either way every chunk is taken by exactly one workgroup.
__device__ __forceinline__ void chunk_iter(int nchunks, int sweep, int& first, int& step,
                                           int& end) {
    const int nblk = gridDim.x;
    if (!sweep) {
        chunk_range(nchunks, first, end);
        step = 1;
    } else if (nblk < 8) {
        first = blockIdx.x;
        step = nblk;
        end = nchunks;
    } else {
        const int x = blockIdx.x & 7, k = blockIdx.x >> 3;
        first = static_cast<int>(static_cast<long long>(x) * nchunks / 8) + k;
        end = static_cast<int>(static_cast<long long>(x + 1) * nchunks / 8);
        step = (nblk - x + 7) >> 3;
    }
}

}  // namespace

// Density terms of G list slots of atom i (slots base, base+TPA, ...; slots at or past kv add
// nothing), added to den0 in slot order.  Knots come from the LDS window s_pb; a pair below the
// window is flagged and its term added after the group from the global table.
template <int G, int MT, int TPA>
__device__ __forceinline__ double density_group(const AtomsDev& a, const ListDev& l,
                                                const TabDev& t, const BoxDev& b, const LtWin& w,
                                                const double* s_pb, const int* __restrict__ li,
                                                int i, int ti, int kv, int base, double2 pxy,
                                                double pz, double den0, int sub = 0,
                                                unsigned long long* mk = nullptr,
                                                const int* jsp = nullptr) {
    constexpr int S = slot_step<TPA>();
    const int kmax = t.ntab2 - 2;
    const int k0 = base / S;  // the lane's slot number of u = 0 (base = sub + k0 * S)
    int js[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int iw = base + u * S;
        js[u] = jsp ? jsp[u] : iw < kv ? li[static_cast<size_t>(iw) * l.stride] : i;
    }
    unsigned fix = 0;
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int j = js[u];
        double2 qxy;
        double qz;
        if constexpr (TPA == 3) {
            double2 h0, h1;
            pair_gather(a.pos, j, sub, h0, h1);
            qxy = h0;
            qz = h1.x;
        } else {
            qxy = *reinterpret_cast<const double2*>(&a.pos[j]);
            qz = a.pos[j].z;
        }
        const double sx = min_image(pxy.x - qxy.x, b.len[0], b.ilen[0]);
        const double sy = min_image(pxy.y - qxy.y, b.len[1], b.ilen[1]);
        const double sz = min_image(pz - qz, b.len[2], b.ilen[2]);
        const double r2 = sx * sx + sy * sy + sz * sz;
        const int row = MT ? t.kpair[ti * t.ngroup + a.ityp[j]] * w.wd : 0;
        double inv_r;
        const double sk = table_coord(r2, t.csi, inv_r);
        const int kk = min(static_cast<int>(sk), kmax);
        const bool in = base + u * S < kv && r2 <= t.ru2max;
        if (mk && in) *mk |= 1ull << (k0 + u);
        const int kl = kk - w.k0d;
        if (kl < 0 && in) fix |= 1u << u;  // below the LDS window: done below
        const int q = row + max(kl, 0);
        const double p0 = s_pb[q], p1 = s_pb[q + 1];
        const double v = p0 + (sk - kk) * (p1 - p0);
        den0 = den0 + ((in && kl >= 0) ? v : 0.0);
    }
    if (fix) {
        // rare close pairs: the same term with the knots of the global table
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (!(fix & (1u << u))) continue;
            const int j = li[static_cast<size_t>(base + u * S) * l.stride];
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
            const double2 e = t.b2[static_cast<size_t>(kind) * t.ntab2 + kk];
            den0 = den0 + (e.x + (sk - kk) * e.y);
        }
    }
    return den0;
}

// Force terms of G list slots of atom i (as density_group), added to (fpx, fpy, fpz) in slot
// order; {FPOTR, FPOTB} knots from the LDS window s_ft.
template <int G, int MT, int TPA>
__device__ __forceinline__ void force_group(const AtomsDev& a, const ListDev& l, const TabDev& t,
                                            const BoxDev& b, const LtWin& w, const double2* s_ft,
                                            const int* __restrict__ li, int i, int ti, int kv,
                                            int base, const double4& pi, double denki,
                                            double& fpx, double& fpy, double& fpz, int sub = 0,
                                            const int* jsp = nullptr) {
    constexpr int S = slot_step<TPA>();
    const int kmax = t.ntab2 - 2;
    int js[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int iw = base + u * S;
        js[u] = jsp ? jsp[u] : iw < kv ? li[static_cast<size_t>(iw) * l.stride] : i;
    }
    unsigned fix = 0;
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int j = js[u];
        double4 pj;
        if constexpr (TPA == 3) {
            double2 h0, h1;
            pair_gather(a.pos, j, sub, h0, h1);
            pj = make_double4(h0.x, h0.y, h1.x, h1.y);
        } else {
            pj = a.pos[j];
        }
        const double sx = min_image(pi.x - pj.x, b.len[0], b.ilen[0]);
        const double sy = min_image(pi.y - pj.y, b.len[1], b.ilen[1]);
        const double sz = min_image(pi.z - pj.z, b.len[2], b.ilen[2]);
        const double r2 = sx * sx + sy * sy + sz * sz;
        int row = 0, row1 = 0;
        if constexpr (MT) {
            const int tj = a.ityp[j];
            row = t.kpair[ti * t.ngroup + tj] * w.wf;
            row1 = t.kpair[tj * t.ngroup + ti] * w.wf;
        }
        double inv_r;
        const double sk = table_coord(r2, t.csi, inv_r);
        const int kk = min(static_cast<int>(sk), kmax);
        const double dk = sk - kk;
        const bool in = base + u * S < kv && r2 <= t.ru2max;
        const int kl = kk - w.k0f;
        if (kl < 0 && in) fix |= 1u << u;  // below the LDS window: done below
        const int q = max(kl, 0);
        const double2 t0 = s_ft[row + q], t1 = s_ft[row + q + 1];
        const double fb = t0.y + dk * (t1.y - t0.y);
        double fb1 = fb;
        if constexpr (MT) {
            const double b0 = s_ft[row1 + q].y, b1 = s_ft[row1 + q + 1].y;
            fb1 = b0 + dk * (b1 - b0);
        }
        // FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r  (CALFORCE_KERNEL)
        const double ft = (t0.x + dk * (t1.x - t0.x)) * (inv_r * inv_r) +
                          (fb * denki + fb1 * pj.w) * inv_r;
        const double fortot = (in && kl >= 0) ? ft : 0.0;
        fpx = fpx + fortot * sx;
        fpy = fpy + fortot * sy;
        fpz = fpz + fortot * sz;
    }
    if (fix) {
        // rare close pairs: the same term with the knots of the global table
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (!(fix & (1u << u))) continue;
            const int j = li[static_cast<size_t>(base + u * S) * l.stride];
            const double4 pj = a.pos[j];
            const double sx = min_image(pi.x - pj.x, b.len[0], b.ilen[0]);
            const double sy = min_image(pi.y - pj.y, b.len[1], b.ilen[1]);
            const double sz = min_image(pi.z - pj.z, b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            int kind = t.kpair[0], kind1 = t.kpair[0];
            if constexpr (MT) {
                const int tj = a.ityp[j];
                kind = t.kpair[ti * t.ngroup + tj];
                kind1 = t.kpair[tj * t.ngroup + ti];
            }
            double inv_r;
            const double sk = table_coord(r2, t.csi, inv_r);
            const int kk = min(static_cast<int>(sk), kmax);
            const double dk = sk - kk;
            const double4 f = t.f4[static_cast<size_t>(kind) * t.ntab2 + kk];
            const double4 f1 = t.f4[static_cast<size_t>(kind1) * t.ntab2 + kk];
            const double fortot = (f.x + dk * f.y) * (inv_r * inv_r) +
                                  ((f.z + dk * f.w) * denki +
                                   (f1.z + dk * f1.w) * pj.w) * inv_r;
            fpx = fpx + fortot * sx;
            fpy = fpy + fortot * sy;
            fpz = fpz + fortot * sz;
        }
    }
}

// Software-pipelined full groups.  js holds the neighbour indices of this group (slots base,
// base + S, ...); the group's position gathers are issued first, then the index loads of the next
// group (slots base + G*S, ..., into js), and only then does the arithmetic wait for the gathers,
// so the next group's index latency hides behind this group's gathers and arithmetic instead of
// adding a second round trip.  Slots at or past kv read the atom itself and add nothing.
// The slot stride of the list is the constant kListRow, so the G loads sit at fixed offsets from
// slot base.  Each load is predicated on its slot being below kv: issuing them all and discarding
// the values measured 10% slower in the density pass, since most slots past an atom's list length
// lie on lines of the list array that no other load touches.
template <int G, int TPA>
__device__ __forceinline__ void load_slots(const int* __restrict__ li, int base, int kv, int self,
                                           int (&js)[G]) {
    constexpr int S = slot_step<TPA>();
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int iw = base + u * S;
        js[u] = iw < kv ? li[static_cast<size_t>(iw) * kListRow] : self;
    }
}

// Record j of the position array (j >= 0: zero-extended index).
__device__ __forceinline__ const double4& rec(const double4* __restrict__ p, int j) {
    return p[static_cast<unsigned>(j)];
}

// Pipelined groups are full: every slot lies below kv.  FW: the LDS window holds the whole table
// (k0 = 0), so no pair falls below it.
template <int G, int MT, int TPA, bool FW>
__device__ __forceinline__ double density_group_pf(const AtomsDev& a, const ListDev& l,
                                                   const TabDev& t, const BoxDev& b,
                                                   const LtWin& w, const double* s_pb,
                                                   const int* __restrict__ li, int i, int ti,
                                                   int kv, int base, double2 pxy, double pz,
                                                   double den0, int (&js)[G],
                                                   unsigned long long* mk = nullptr) {
    constexpr int S = slot_step<TPA>();
    const int kmax = t.ntab2 - 2;
    const int k0 = base / S;
    double2 qxy[G];
    double qz[G];
    int tj[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
        qxy[u] = *reinterpret_cast<const double2*>(&rec(a.pos, js[u]));
        qz[u] = rec(a.pos, js[u]).z;
        tj[u] = MT ? a.ityp[js[u]] : 0;
    }
    load_slots<G, TPA>(li, base + G * S, kv, i, js);
    unsigned fix = 0;
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const double sx = min_image(pxy.x - qxy[u].x, b.len[0], b.ilen[0]);
        const double sy = min_image(pxy.y - qxy[u].y, b.len[1], b.ilen[1]);
        const double sz = min_image(pz - qz[u], b.len[2], b.ilen[2]);
        const double r2 = sx * sx + sy * sy + sz * sz;
        const int row = MT ? t.kpair[ti * t.ngroup + tj[u]] * w.wd : 0;
        double inv_r;
        const double sk = table_coord(r2, t.csi, inv_r);
        const int kk = min(static_cast<int>(sk), kmax);
        const bool in = r2 <= t.ru2max;
        if (mk && in) *mk |= 1ull << (k0 + u);
        const int kl = FW ? kk : kk - w.k0d;
        if (!FW && kl < 0 && in) fix |= 1u << u;  // below the LDS window: done below
        const int q = row + (FW ? kl : max(kl, 0));
        const double p0 = s_pb[q], p1 = s_pb[q + 1];
        const double v = p0 + (sk - kk) * (p1 - p0);
        den0 = den0 + ((in && (FW || kl >= 0)) ? v : 0.0);
    }
    if (!FW && fix) {
        // rare close pairs: reload the neighbour (keeping the group's records live would cost
        // registers) and take the knots of the global table
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (!(fix & (1u << u))) continue;
            const int j = li[static_cast<size_t>(base + u * S) * l.stride];
            const double2 q2 = *reinterpret_cast<const double2*>(&a.pos[j]);
            const double sx = min_image(pxy.x - q2.x, b.len[0], b.ilen[0]);
            const double sy = min_image(pxy.y - q2.y, b.len[1], b.ilen[1]);
            const double sz = min_image(pz - a.pos[j].z, b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            const int kind = MT ? t.kpair[ti * t.ngroup + a.ityp[j]] : t.kpair[0];
            double inv_r;
            const double sk = table_coord(r2, t.csi, inv_r);
            const int kk = min(static_cast<int>(sk), kmax);
            const double2 e = t.b2[static_cast<size_t>(kind) * t.ntab2 + kk];
            den0 = den0 + (e.x + (sk - kk) * e.y);
        }
    }
    return den0;
}

// As density_group_pf: full groups; FW: the window holds the whole table.
template <int G, int MT, int TPA, bool FW>
__device__ __forceinline__ void force_group_pf(const AtomsDev& a, const ListDev& l,
                                               const TabDev& t, const BoxDev& b, const LtWin& w,
                                               const double2* s_ft, const int* __restrict__ li,
                                               int i, int ti, int kv, int base, const double4& pi,
                                               double denki, double& fpx, double& fpy,
                                               double& fpz, int (&js)[G]) {
    constexpr int S = slot_step<TPA>();
    const int kmax = t.ntab2 - 2;
    double4 pj[G];
    int tj[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
        pj[u] = rec(a.pos, js[u]);
        tj[u] = MT ? a.ityp[js[u]] : 0;
    }
    load_slots<G, TPA>(li, base + G * S, kv, i, js);
    unsigned fix = 0;
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const double sx = min_image(pi.x - pj[u].x, b.len[0], b.ilen[0]);
        const double sy = min_image(pi.y - pj[u].y, b.len[1], b.ilen[1]);
        const double sz = min_image(pi.z - pj[u].z, b.len[2], b.ilen[2]);
        const double r2 = sx * sx + sy * sy + sz * sz;
        int row = 0, row1 = 0;
        if constexpr (MT) {
            row = t.kpair[ti * t.ngroup + tj[u]] * w.wf;
            row1 = t.kpair[tj[u] * t.ngroup + ti] * w.wf;
        }
        double inv_r;
        const double sk = table_coord(r2, t.csi, inv_r);
        const int kk = min(static_cast<int>(sk), kmax);
        const double dk = sk - kk;
        const bool in = r2 <= t.ru2max;
        const int kl = FW ? kk : kk - w.k0f;
        if (!FW && kl < 0 && in) fix |= 1u << u;  // below the LDS window: done below
        const int q = FW ? kl : max(kl, 0);
        const double2 t0 = s_ft[row + q], t1 = s_ft[row + q + 1];
        const double fb = t0.y + dk * (t1.y - t0.y);
        double fb1 = fb;
        if constexpr (MT) {
            const double b0 = s_ft[row1 + q].y, b1 = s_ft[row1 + q + 1].y;
            fb1 = b0 + dk * (b1 - b0);
        }
        // FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r  (CALFORCE_KERNEL)
        const double ft = (t0.x + dk * (t1.x - t0.x)) * (inv_r * inv_r) +
                          (fb * denki + fb1 * pj[u].w) * inv_r;
        const double fortot = (in && (FW || kl >= 0)) ? ft : 0.0;
        fpx = fpx + fortot * sx;
        fpy = fpy + fortot * sy;
        fpz = fpz + fortot * sz;
    }
    if (!FW && fix) {
        // rare close pairs: reload the neighbour and take the knots of the global table
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (!(fix & (1u << u))) continue;
            const int j = li[static_cast<size_t>(base + u * S) * l.stride];
            const double4 pk = a.pos[j];
            const double sx = min_image(pi.x - pk.x, b.len[0], b.ilen[0]);
            const double sy = min_image(pi.y - pk.y, b.len[1], b.ilen[1]);
            const double sz = min_image(pi.z - pk.z, b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            int kind = t.kpair[0], kind1 = t.kpair[0];
            if constexpr (MT) {
                const int tk = a.ityp[j];
                kind = t.kpair[ti * t.ngroup + tk];
                kind1 = t.kpair[tk * t.ngroup + ti];
            }
            double inv_r;
            const double sk = table_coord(r2, t.csi, inv_r);
            const int kk = min(static_cast<int>(sk), kmax);
            const double dk = sk - kk;
            const double4 f = t.f4[static_cast<size_t>(kind) * t.ntab2 + kk];
            const double4 f1 = t.f4[static_cast<size_t>(kind1) * t.ntab2 + kk];
            const double fortot = (f.x + dk * f.y) * (inv_r * inv_r) +
                                  ((f.z + dk * f.w) * denki + (f1.z + dk * f1.w) * pk.w) * inv_r;
            fpx = fpx + fortot * sx;
            fpy = fpy + fortot * sy;
            fpz = fpz + fortot * sz;
        }
    }
}

// Force terms over the slots the density pass found inside the cutoff: m holds the lane's
// in-cutoff slot numbers as bits (bit k = slot sub + k * S); groups of G of them, software
// pipelined as force_group_pf.  The pairs skipped are those whose term is zero, so the sums are
// those of the full walk.
template <int G, int TPA>
__device__ __forceinline__ void pop_slots(const int* __restrict__ li, int stride, int sub,
                                          unsigned long long& m, int self, int (&js)[G],
                                          int (&ks)[G]) {
    constexpr int S = slot_step<TPA>();
#pragma unroll
    for (int u = 0; u < G; ++u) {
        ks[u] = -1;
        js[u] = self;
        if (m) {
            const int k = __builtin_ctzll(m);
            m &= m - 1;
            ks[u] = sub + k * S;
            js[u] = li[static_cast<size_t>(ks[u]) * stride];
        }
    }
}

template <int G, int MT, int TPA>
__device__ __forceinline__ void force_masked(const AtomsDev& a, const ListDev& l, const TabDev& t,
                                             const BoxDev& b, const LtWin& w,
                                             const double2* s_ft, const int* __restrict__ li,
                                             int i, int ti, int sub, unsigned long long m,
                                             const double4& pi, double denki, double& fpx,
                                             double& fpy, double& fpz) {
    const int kmax = t.ntab2 - 2;
    int js[G], ks[G];
    pop_slots<G, TPA>(li, l.stride, sub, m, i, js, ks);
    while (ks[0] >= 0) {
        double4 pj[G];
        int tj[G], kc[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            pj[u] = a.pos[js[u]];
            tj[u] = MT ? a.ityp[js[u]] : 0;
            kc[u] = ks[u];
        }
        pop_slots<G, TPA>(li, l.stride, sub, m, i, js, ks);
        unsigned fix = 0;
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const double sx = min_image(pi.x - pj[u].x, b.len[0], b.ilen[0]);
            const double sy = min_image(pi.y - pj[u].y, b.len[1], b.ilen[1]);
            const double sz = min_image(pi.z - pj[u].z, b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            int row = 0, row1 = 0;
            if constexpr (MT) {
                row = t.kpair[ti * t.ngroup + tj[u]] * w.wf;
                row1 = t.kpair[tj[u] * t.ngroup + ti] * w.wf;
            }
            double inv_r;
            const double sk = table_coord(r2, t.csi, inv_r);
            const int kk = min(static_cast<int>(sk), kmax);
            const double dk = sk - kk;
            const bool in = kc[u] >= 0 && r2 <= t.ru2max;
            const int kl = kk - w.k0f;
            if (kl < 0 && in) fix |= 1u << u;  // below the LDS window: done below
            const int q = max(kl, 0);
            const double2 t0 = s_ft[row + q], t1 = s_ft[row + q + 1];
            const double fb = t0.y + dk * (t1.y - t0.y);
            double fb1 = fb;
            if constexpr (MT) {
                const double b0 = s_ft[row1 + q].y, b1 = s_ft[row1 + q + 1].y;
                fb1 = b0 + dk * (b1 - b0);
            }
            // FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r  (CALFORCE_KERNEL)
            const double ft = (t0.x + dk * (t1.x - t0.x)) * (inv_r * inv_r) +
                              (fb * denki + fb1 * pj[u].w) * inv_r;
            const double fortot = (in && kl >= 0) ? ft : 0.0;
            fpx = fpx + fortot * sx;
            fpy = fpy + fortot * sy;
            fpz = fpz + fortot * sz;
        }
        if (fix) {
            // rare close pairs: reload the neighbour and take the knots of the global table
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (!(fix & (1u << u))) continue;
                const int j = li[static_cast<size_t>(kc[u]) * l.stride];
                const double4 pk = a.pos[j];
                const double sx = min_image(pi.x - pk.x, b.len[0], b.ilen[0]);
                const double sy = min_image(pi.y - pk.y, b.len[1], b.ilen[1]);
                const double sz = min_image(pi.z - pk.z, b.len[2], b.ilen[2]);
                const double r2 = sx * sx + sy * sy + sz * sz;
                int kind = t.kpair[0], kind1 = t.kpair[0];
                if constexpr (MT) {
                    const int tk = a.ityp[j];
                    kind = t.kpair[ti * t.ngroup + tk];
                    kind1 = t.kpair[tk * t.ngroup + ti];
                }
                double inv_r;
                const double sk = table_coord(r2, t.csi, inv_r);
                const int kk = min(static_cast<int>(sk), kmax);
                const double dk = sk - kk;
                const double4 f = t.f4[static_cast<size_t>(kind) * t.ntab2 + kk];
                const double4 f1 = t.f4[static_cast<size_t>(kind1) * t.ntab2 + kk];
                const double fortot =
                    (f.x + dk * f.y) * (inv_r * inv_r) +
                    ((f.z + dk * f.w) * denki + (f1.z + dk * f1.w) * pk.w) * inv_r;
                fpx = fpx + fortot * sx;
                fpy = fpy + fortot * sy;
                fpz = fpz + fortot * sz;
            }
        }
    }
}

// Knot rows in LDS: kind-major, w knots per kind from knot k0 (all kinds of the table for
// several atom types, the one kind in use for a single type).
template <int MB, int MT, int EAM, int TPA>
__global__ __launch_bounds__(kLtMaxBlock) void k_density_lt(AtomsDev a, ListDev l, TabDev t,
                                                            BoxDev box0,
                                                            const BoxDev* __restrict__ boxes,
                                                            LtWin w, int nchunks, int* err) {
    extern __shared__ double s_pb[];
    const int nk = MT ? t.nkind : 1;
    for (int k = threadIdx.x; k < nk * w.wd; k += blockDim.x) {
        const int kd = MT ? k / w.wd : t.kpair[0], q = MT ? k % w.wd : k;
        s_pb[k] = t.b2[static_cast<size_t>(kd) * t.ntab2 + w.k0d + q].x;
    }
    __syncthreads();
    int c, cstep, cend;
    chunk_iter(nchunks, w.sweep, c, cstep, cend);
    const int kmax = t.ntab2 - 2;
    for (; c < cend; c += cstep) {
        int i, sub;
        atom_of_lane<TPA>(c, i, sub);
        const bool have = i < a.n;
        // The atom's loads in one batch, issued before its status is known: status, position,
        // list length and, on the plain list walk (TPA 1, 2, 4), the first group's slot indices
        // (slots clamped below mx, masked by the length once it is known), so that the first
        // gathers wait on one round trip instead of three.
        int st = 0, kv = 0;
        double2 pxy = make_double2(0.0, 0.0);
        double pz = 0.0;
        int js0[kGroupD];
        const int* __restrict__ li = l.indi + list_base(have ? i : 0, l.mx);
        if (have) {
            st = a.statu[i];
            pxy = *reinterpret_cast<const double2*>(&a.pos[i]);
            pz = a.pos[i].z;
            kv = l.kvois[i];
            if constexpr (TPA != 3) {
                constexpr int S = slot_step<TPA>();
#pragma unroll
                for (int u = 0; u < kGroupD; ++u)
                    js0[u] = li[static_cast<size_t>(min(sub + u * S, l.mx - 1)) * kListRow];
            }
        }
        const bool active = have && (st & STATU_ACTIVE) == STATU_ACTIVE;
        double den0 = 0.0;
        int ti = 0;
        if (active) {
            const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, i);
            ti = MT ? a.ityp[i] : 0;
            // the slots of a lane are visited in increasing order on every path
            if constexpr (TPA == 3) {
                // the two lanes of a pair run the same groups (their gathers are shared): full
                // groups of kGroup slots, then groups of 4 and single slots for the tail
                int base = 0;
                for (; base + 2 * kGroup <= kv; base += 2 * kGroup)
                    den0 = density_group<kGroup, MT, 3>(a, l, t, b, w, s_pb, li, i, ti, kv,
                                                        base + sub, pxy, pz, den0, sub);
                for (; base + 8 <= kv; base += 8)
                    den0 = density_group<4, MT, 3>(a, l, t, b, w, s_pb, li, i, ti, kv, base + sub,
                                                   pxy, pz, den0, sub);
                for (; base < kv; base += 2)
                    den0 = density_group<1, MT, 3>(a, l, t, b, w, s_pb, li, i, ti, kv, base + sub,
                                                   pxy, pz, den0, sub);
            } else {
                // two copies of the slot loop: the in-cutoff slot mask is built only when the
                // force pass reads it (a run-time flag would keep its bit arithmetic in the loop)
                auto slots = [&](auto with_mask) {
                    constexpr bool WM = decltype(with_mask)::value;
                    int base = sub;
                    int js[kGroupD];
                    unsigned long long mk = 0;  // this lane's in-cutoff slots
                    unsigned long long* mkp = WM ? &mk : nullptr;
#pragma unroll
                    for (int u = 0; u < kGroupD; ++u) js[u] = base + u * TPA < kv ? js0[u] : i;
                    auto full = [&](auto fw) {
                        constexpr bool FW = decltype(fw)::value;
                        for (; base + (kGroupD - 1) * TPA < kv; base += TPA * kGroupD)
                            den0 = density_group_pf<kGroupD, MT, TPA, FW>(
                                a, l, t, b, w, s_pb, li, i, ti, kv, base, pxy, pz, den0, js, mkp);
                    };
                    if (w.k0d == 0)
                        full(std::true_type{});
                    else
                        full(std::false_type{});
                    // the last kGroupD - 1 slots at most, in pairs and singles, from the indices
                    // the last group (or the atom's batch) already read into js
#pragma unroll
                    for (int u = 0; u < kGroupD - 1; u += 2) {
                        const int iw = base + u * TPA;
                        if (iw + TPA < kv)
                            den0 = density_group<2, MT, TPA>(a, l, t, b, w, s_pb, li, i, ti, kv,
                                                             iw, pxy, pz, den0, sub, mkp, &js[u]);
                        else if (iw < kv)
                            den0 = density_group<1, MT, TPA>(a, l, t, b, w, s_pb, li, i, ti, kv,
                                                             iw, pxy, pz, den0, sub, mkp, &js[u]);
                    }
                    if constexpr (WM) l.cmask[static_cast<size_t>(i) * TPA + sub] = mk;
                };
                if (l.cmask)
                    slots(std::true_type{});
                else
                    slots(std::false_type{});
            }
        }
        den0 = lane_sum<TPA>(den0);
        if (sub != 0 || !have) continue;
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
}

template <int MB, int MT, int MODE, int TPA>
__global__ __launch_bounds__(kLtMaxBlock) void k_force_lt(AtomsDev a, ListDev l, TabDev t,
                                                          BoxDev box0,
                                                          const BoxDev* __restrict__ boxes,
                                                          StepParams sp, LtWin w, int nchunks) {
    extern __shared__ double2 s_ft[];
    const int nk = MT ? t.nkind : 1;
    for (int k = threadIdx.x; k < nk * w.wf; k += blockDim.x) {
        const int kd = MT ? k / w.wf : t.kpair[0], q = MT ? k % w.wf : k;
        const double4 f = t.f4[static_cast<size_t>(kd) * t.ntab2 + w.k0f + q];
        s_ft[k] = make_double2(f.x, f.z);
    }
    __syncthreads();
    int c, cstep, cend;
    chunk_iter(nchunks, w.sweep, c, cstep, cend);
    const int kmax = t.ntab2 - 2;
    for (; c < cend; c += cstep) {
        int i, sub;
        atom_of_lane<TPA>(c, i, sub);
        const bool have = i < a.n;
        const int st = have ? a.statu[i] : 0;
        const bool active = have && (st & STATU_ACTIVE) == STATU_ACTIVE;
        const double4 pi = have ? a.pos[i] : make_double4(0, 0, 0, 0);
        const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, have ? i : 0);
        const int ti = (MT && have) ? a.ityp[i] : 0;
        double fpx = 0.0, fpy = 0.0, fpz = 0.0;
        // The atom's loads in one batch, issued before its status is known: the integrator's
        // inputs (by the atom's first lane), the list length and, on the plain list walk, the
        // first group's slot indices (slots clamped below mx, masked by the length once it is
        // known), so that the first gathers wait on one round trip instead of three.
        AtomState as;
        int kv = 0;
        int js[kGroupF];
        const bool walk = TPA != 3 && !l.cmask;
        const int* __restrict__ li = l.indi + list_base(have ? i : 0, l.mx);
        if (have) {
            if constexpr (MODE >= 1)
                if (sub == 0) load_atom_state(a, sp, i, as);
            kv = l.kvois[i];
            if (walk) {
                constexpr int S = slot_step<TPA>();
#pragma unroll
                for (int u = 0; u < kGroupF; ++u)
                    js[u] = li[static_cast<size_t>(min(sub + u * S, l.mx - 1)) * kListRow];
            }
        }
        if (active) {
            const double denki = pi.w;
            if constexpr (TPA == 3) {
                // the two lanes of a pair run the same groups (their gathers are shared)
                int base = 0;
                for (; base + 2 * kGroup <= kv; base += 2 * kGroup)
                    force_group<kGroup, MT, 3>(a, l, t, b, w, s_ft, li, i, ti, kv, base + sub, pi,
                                               denki, fpx, fpy, fpz, sub);
                for (; base + 8 <= kv; base += 8)
                    force_group<4, MT, 3>(a, l, t, b, w, s_ft, li, i, ti, kv, base + sub, pi, denki,
                                          fpx, fpy, fpz, sub);
                for (; base < kv; base += 2)
                    force_group<1, MT, 3>(a, l, t, b, w, s_ft, li, i, ti, kv, base + sub, pi, denki,
                                          fpx, fpy, fpz, sub);
            } else if (l.cmask) {
                force_masked<kGroupF, MT, TPA>(a, l, t, b, w, s_ft, li, i, ti, sub,
                                               l.cmask[static_cast<size_t>(i) * TPA + sub], pi,
                                               denki, fpx, fpy, fpz);
            } else {
                int base = sub;
#pragma unroll
                for (int u = 0; u < kGroupF; ++u)
                    if (base + u * TPA >= kv) js[u] = i;
                auto full = [&](auto fw) {
                    constexpr bool FW = decltype(fw)::value;
                    for (; base + (kGroupF - 1) * TPA < kv; base += TPA * kGroupF)
                        force_group_pf<kGroupF, MT, TPA, FW>(a, l, t, b, w, s_ft, li, i, ti, kv,
                                                             base, pi, denki, fpx, fpy, fpz, js);
                };
                if (w.k0f == 0)
                    full(std::true_type{});
                else
                    full(std::false_type{});
                // the last kGroupF - 1 slots at most, in pairs and singles, from the indices the
                // last group (or the atom's batch) already read into js
#pragma unroll
                for (int u = 0; u < kGroupF - 1; u += 2) {
                    const int iw = base + u * TPA;
                    if (iw + TPA < kv)
                        force_group<2, MT, TPA>(a, l, t, b, w, s_ft, li, i, ti, kv, iw, pi, denki,
                                                fpx, fpy, fpz, sub, &js[u]);
                    else if (iw < kv)
                        force_group<1, MT, TPA>(a, l, t, b, w, s_ft, li, i, ti, kv, iw, pi, denki,
                                                fpx, fpy, fpz, sub, &js[u]);
                }
            }
        }
        fpx = lane_sum<TPA>(fpx);
        fpy = lane_sum<TPA>(fpy);
        fpz = lane_sum<TPA>(fpz);
        if (sub != 0 || !have) continue;
        if (sp.write_f) {
            a.fx[i] = fpx;
            a.fy[i] = fpy;
            a.fz[i] = fpz;
        }
        if constexpr (MODE >= 1)
            integrate_atom<MODE>(a, t, b, sp, i, st, active, pi, ti, fpx, fpy, fpz, &as);
    }
}

// --------------------------------------------------------------------------------------------
// LDS bricks with the table window in LDS
//
// Each workgroup holds a table window (as above) and walks a contiguous range of bricks (those of
// one XCD are contiguous, so neighbouring bricks share halo cells in that XCD's L2).  Per brick it
// stages the halo atoms in LDS with coalesced loads and then every home atom reads its list as
// 16-bit halo slots: positions and knots both come from LDS, and the only per-pair global load is
// the 2-byte slot.  A brick whose halo exceeds the staging capacity `cap` takes the global-list groups
// above (same results).  The list must have been built by the brick list kernels on the same
// brick plan (the slots are positions in that brick's halo).
// --------------------------------------------------------------------------------------------

// Density terms of G list slots of the home atom at halo slot si (as density_group, with the
// neighbour positions read from the staged halo x/y/z arrays).
template <int G, int MT, int TPA>
__device__ __forceinline__ double density_group_s(const TabDev& t, const BoxDev& b,
                                                  const LtWin& w, const double* s_pb,
                                                  const double* lx, const double* ly,
                                                  const double* lz, const int* lty,
                                                  const unsigned short* __restrict__ li,
                                                  int stride, int si, int ti, int kv, int base,
                                                  double xi, double yi, double zi, double den0) {
    const int kmax = t.ntab2 - 2;
    int qs[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int iw = base + u * TPA;
        qs[u] = iw < kv ? li[static_cast<size_t>(iw) * stride] : si;
    }
    unsigned fix = 0;
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int q = qs[u];
        const double sx = min_image(xi - lx[q], b.len[0], b.ilen[0]);
        const double sy = min_image(yi - ly[q], b.len[1], b.ilen[1]);
        const double sz = min_image(zi - lz[q], b.len[2], b.ilen[2]);
        const double r2 = sx * sx + sy * sy + sz * sz;
        const int row = MT ? t.kpair[ti * t.ngroup + lty[q]] * w.wd : 0;
        double inv_r;
        const double sk = table_coord(r2, t.csi, inv_r);
        const int kk = min(static_cast<int>(sk), kmax);
        const bool in = base + u * TPA < kv && r2 <= t.ru2max;
        const int kl = kk - w.k0d;
        if (kl < 0 && in) fix |= 1u << u;  // below the LDS window: done below
        const int k2 = row + max(kl, 0);
        const double p0 = s_pb[k2], p1 = s_pb[k2 + 1];
        const double v = p0 + (sk - kk) * (p1 - p0);
        den0 = den0 + ((in && kl >= 0) ? v : 0.0);
    }
    if (fix) {
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (!(fix & (1u << u))) continue;
            const int q = qs[u];
            const double sx = min_image(xi - lx[q], b.len[0], b.ilen[0]);
            const double sy = min_image(yi - ly[q], b.len[1], b.ilen[1]);
            const double sz = min_image(zi - lz[q], b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            const int kind = MT ? t.kpair[ti * t.ngroup + lty[q]] : t.kpair[0];
            double inv_r;
            const double sk = table_coord(r2, t.csi, inv_r);
            const int kk = min(static_cast<int>(sk), kmax);
            const double2 e = t.b2[static_cast<size_t>(kind) * t.ntab2 + kk];
            den0 = den0 + (e.x + (sk - kk) * e.y);
        }
    }
    return den0;
}

// Force terms of G list slots of the home atom at halo slot si (as force_group, with the
// neighbour positions and F' read from the staged halo).
template <int G, int MT, int TPA>
__device__ __forceinline__ void force_group_s(const TabDev& t, const BoxDev& b, const LtWin& w,
                                              const double2* s_ft, const double4* lp,
                                              const int* lty,
                                              const unsigned short* __restrict__ li, int stride,
                                              int si, int ti, int kv, int base, const double4& pi,
                                              double& fpx, double& fpy, double& fpz) {
    const int kmax = t.ntab2 - 2;
    int qs[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int iw = base + u * TPA;
        qs[u] = iw < kv ? li[static_cast<size_t>(iw) * stride] : si;
    }
    unsigned fix = 0;
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int q = qs[u];
        const double4 pj = lp[q];
        const double sx = min_image(pi.x - pj.x, b.len[0], b.ilen[0]);
        const double sy = min_image(pi.y - pj.y, b.len[1], b.ilen[1]);
        const double sz = min_image(pi.z - pj.z, b.len[2], b.ilen[2]);
        const double r2 = sx * sx + sy * sy + sz * sz;
        int row = 0, row1 = 0;
        if constexpr (MT) {
            const int tj = lty[q];
            row = t.kpair[ti * t.ngroup + tj] * w.wf;
            row1 = t.kpair[tj * t.ngroup + ti] * w.wf;
        }
        double inv_r;
        const double sk = table_coord(r2, t.csi, inv_r);
        const int kk = min(static_cast<int>(sk), kmax);
        const double dk = sk - kk;
        const bool in = base + u * TPA < kv && r2 <= t.ru2max;
        const int kl = kk - w.k0f;
        if (kl < 0 && in) fix |= 1u << u;  // below the LDS window: done below
        const int k2 = max(kl, 0);
        const double2 t0 = s_ft[row + k2], t1 = s_ft[row + k2 + 1];
        const double fb = t0.y + dk * (t1.y - t0.y);
        double fb1 = fb;
        if constexpr (MT) {
            const double b0 = s_ft[row1 + k2].y, b1 = s_ft[row1 + k2 + 1].y;
            fb1 = b0 + dk * (b1 - b0);
        }
        // FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r  (CALFORCE_KERNEL)
        const double ft = (t0.x + dk * (t1.x - t0.x)) * (inv_r * inv_r) +
                          (fb * pi.w + fb1 * pj.w) * inv_r;
        const double fortot = (in && kl >= 0) ? ft : 0.0;
        fpx = fpx + fortot * sx;
        fpy = fpy + fortot * sy;
        fpz = fpz + fortot * sz;
    }
    if (fix) {
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (!(fix & (1u << u))) continue;
            const double4 pj = lp[qs[u]];
            const double sx = min_image(pi.x - pj.x, b.len[0], b.ilen[0]);
            const double sy = min_image(pi.y - pj.y, b.len[1], b.ilen[1]);
            const double sz = min_image(pi.z - pj.z, b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            int kind = t.kpair[0], kind1 = t.kpair[0];
            if constexpr (MT) {
                const int tj = lty[qs[u]];
                kind = t.kpair[ti * t.ngroup + tj];
                kind1 = t.kpair[tj * t.ngroup + ti];
            }
            double inv_r;
            const double sk = table_coord(r2, t.csi, inv_r);
            const int kk = min(static_cast<int>(sk), kmax);
            const double dk = sk - kk;
            const double4 f = t.f4[static_cast<size_t>(kind) * t.ntab2 + kk];
            const double4 f1 = t.f4[static_cast<size_t>(kind1) * t.ntab2 + kk];
            const double fortot = (f.x + dk * f.y) * (inv_r * inv_r) +
                                  ((f.z + dk * f.w) * pi.w + (f1.z + dk * f1.w) * pj.w) * inv_r;
            fpx = fpx + fortot * sx;
            fpy = fpy + fortot * sy;
            fpz = fpz + fortot * sz;
        }
    }
}

// Slots per group of the staged loops (LDS latency is short; larger groups spill).
constexpr int kGroupS = 4;

// Dynamic LDS of the brick passes: the knot window, then the staged halo of at most cap atoms
// (x, y, z arrays for the density pass, double4 records for the force pass; types if MT).
constexpr long bt_atom_bytes(bool force, bool mt) {
    return static_cast<long>((force ? sizeof(double4) : 3 * sizeof(double)) +
                             (mt ? sizeof(int) : 0));
}

template <int MT, int EAM, int TPA>
__global__ __launch_bounds__(kLtMaxBlock) void k_density_bt(const int* __restrict__ brick_box,
                                                            int nbricks, AtomsDev a, ListDev l,
                                                            TabDev t,
                                                            const BoxDev* __restrict__ boxes,
                                                            const int* __restrict__ start,
                                                            LtWin w, int cap, int* err) {
    extern __shared__ double s_dyn[];
    __shared__ BrickLds L;
    const int nk = MT ? t.nkind : 1;
    double* s_pb = s_dyn;
    const int nwin = (nk * w.wd + 1) & ~1;  // keep the halo arrays 16-byte aligned
    double* lx = s_dyn + nwin;
    double* ly = lx + cap;
    double* lz = ly + cap;
    int* lty = reinterpret_cast<int*>(lz + cap);
    for (int k = threadIdx.x; k < nk * w.wd; k += blockDim.x) {
        const int kd = MT ? k / w.wd : t.kpair[0], q = MT ? k % w.wd : k;
        s_pb[k] = t.b2[static_cast<size_t>(kd) * t.ntab2 + w.k0d + q].x;
    }
    int b0, b1;
    chunk_range(nbricks, b0, b1);
    for (int bi = b0; bi < b1; ++bi) {
        __syncthreads();  // the previous brick's halo is no longer read
        const BoxDev& B = boxes[brick_box[bi]];
        BrickInfo g;
        brick_info(B, bi, g);
        const int total = brick_halo(B, g, start, L.hcell, nullptr, L.hoff, L.homeoff, L.scan);
        const bool staged = total <= cap;
        if (staged) {
            for_halo_slots(g, start, L.hcell, L.hoff, [&](int s, int q, int) {
                // x, y, z only: w (F') of other atoms is being written by this pass
                const double2 xy = *reinterpret_cast<const double2*>(&a.pos[q]);
                lx[s] = xy.x;
                ly[s] = xy.y;
                lz[s] = a.pos[q].z;
                if constexpr (MT) lty[s] = a.ityp[q];
            });
        }
        __syncthreads();
        const int nlanes = L.homeoff[g.nhc] * TPA;
        for (int t0 = 0; t0 < nlanes; t0 += blockDim.x) {
            const int tt = t0 + threadIdx.x;
            const int ha = tt / TPA, sub = tt % TPA;
            const bool have = tt < nlanes;
            int p = 0, ti = 0;
            bool active = false;
            double den0 = 0.0;
            if (have) {
                int h, off;
                home_atom(ha, g, L.homeoff, h, off);
                p = start[L.hcell[h]] + off;
                active = (a.statu[p] & STATU_ACTIVE) == STATU_ACTIVE;
                if (active) {
                    ti = MT ? a.ityp[p] : 0;
                    const int kv = l.kvois[p];
                    int base = sub;
                    if (staged) {
                        const int si = L.hoff[h] + off;
                        const double xi = lx[si], yi = ly[si], zi = lz[si];
                        const unsigned short* __restrict__ li = l.sidx + list_base(p, l.mx);
                        for (; base + (kGroupS - 1) * TPA < kv; base += TPA * kGroupS)
                            den0 = density_group_s<kGroupS, MT, TPA>(t, B, w, s_pb, lx, ly, lz,
                                                                     lty, li, l.stride, si, ti, kv,
                                                                     base, xi, yi, zi, den0);
                        for (; base < kv; base += TPA)
                            den0 = density_group_s<1, MT, TPA>(t, B, w, s_pb, lx, ly, lz, lty, li,
                                                               l.stride, si, ti, kv, base, xi, yi,
                                                               zi, den0);
                    } else {
                        const double2 pxy = *reinterpret_cast<const double2*>(&a.pos[p]);
                        const double pz = a.pos[p].z;
                        const int* __restrict__ li = l.indi + list_base(p, l.mx);
                        for (; base + 3 * TPA < kv; base += TPA * 4)
                            den0 = density_group<4, MT, TPA>(a, l, t, B, w, s_pb, li, p, ti, kv,
                                                             base, pxy, pz, den0);
                        for (; base < kv; base += TPA)
                            den0 = density_group<1, MT, TPA>(a, l, t, B, w, s_pb, li, p, ti, kv,
                                                             base, pxy, pz, den0);
                    }
                }
            }
            den0 = lane_sum<TPA>(den0);
            if (sub != 0 || !have) continue;
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
            a.pos[p].w = den0;
        }
    }
}

template <int MT, int MODE, int TPA>
__global__ __launch_bounds__(kLtMaxBlock) void k_force_bt(const int* __restrict__ brick_box,
                                                          int nbricks, AtomsDev a, ListDev l,
                                                          TabDev t,
                                                          const BoxDev* __restrict__ boxes,
                                                          const int* __restrict__ start,
                                                          StepParams sp, LtWin w, int cap) {
    extern __shared__ double2 s_dyn2[];
    __shared__ BrickLds L;
    const int nk = MT ? t.nkind : 1;
    double2* s_ft = s_dyn2;
    double4* lp = reinterpret_cast<double4*>(s_dyn2 + ((nk * w.wf + 1) & ~1));
    int* lty = reinterpret_cast<int*>(lp + cap);
    for (int k = threadIdx.x; k < nk * w.wf; k += blockDim.x) {
        const int kd = MT ? k / w.wf : t.kpair[0], q = MT ? k % w.wf : k;
        const double4 f = t.f4[static_cast<size_t>(kd) * t.ntab2 + w.k0f + q];
        s_ft[k] = make_double2(f.x, f.z);
    }
    int b0, b1;
    chunk_range(nbricks, b0, b1);
    for (int bi = b0; bi < b1; ++bi) {
        __syncthreads();  // the previous brick's halo is no longer read
        const BoxDev& B = boxes[brick_box[bi]];
        BrickInfo g;
        brick_info(B, bi, g);
        const int total = brick_halo(B, g, start, L.hcell, nullptr, L.hoff, L.homeoff, L.scan);
        const bool staged = total <= cap;
        if (staged) {
            for_halo_slots(g, start, L.hcell, L.hoff, [&](int s, int q, int) {
                lp[s] = a.pos[q];  // x, y, z and F'
                if constexpr (MT) lty[s] = a.ityp[q];
            });
        }
        __syncthreads();
        const int nlanes = L.homeoff[g.nhc] * TPA;
        for (int t0 = 0; t0 < nlanes; t0 += blockDim.x) {
            const int tt = t0 + threadIdx.x;
            const int ha = tt / TPA, sub = tt % TPA;
            const bool have = tt < nlanes;
            int p = 0, ti = 0, st = 0;
            bool active = false;
            double4 pi = make_double4(0, 0, 0, 0);
            double fpx = 0.0, fpy = 0.0, fpz = 0.0;
            if (have) {
                int h, off;
                home_atom(ha, g, L.homeoff, h, off);
                p = start[L.hcell[h]] + off;
                const int si = L.hoff[h] + off;
                st = a.statu[p];
                active = (st & STATU_ACTIVE) == STATU_ACTIVE;
                pi = staged ? lp[si] : a.pos[p];
                ti = MT ? a.ityp[p] : 0;
                if (active) {
                    const int kv = l.kvois[p];
                    int base = sub;
                    if (staged) {
                        const unsigned short* __restrict__ li = l.sidx + list_base(p, l.mx);
                        for (; base + (kGroupS - 1) * TPA < kv; base += TPA * kGroupS)
                            force_group_s<kGroupS, MT, TPA>(t, B, w, s_ft, lp, lty, li, l.stride,
                                                            si, ti, kv, base, pi, fpx, fpy, fpz);
                        for (; base < kv; base += TPA)
                            force_group_s<1, MT, TPA>(t, B, w, s_ft, lp, lty, li, l.stride, si, ti,
                                                      kv, base, pi, fpx, fpy, fpz);
                    } else {
                        const int* __restrict__ li = l.indi + list_base(p, l.mx);
                        for (; base + 3 * TPA < kv; base += TPA * 4)
                            force_group<4, MT, TPA>(a, l, t, B, w, s_ft, li, p, ti, kv, base, pi,
                                                    pi.w, fpx, fpy, fpz);
                        for (; base < kv; base += TPA)
                            force_group<1, MT, TPA>(a, l, t, B, w, s_ft, li, p, ti, kv, base, pi,
                                                    pi.w, fpx, fpy, fpz);
                    }
                }
            }
            fpx = lane_sum<TPA>(fpx);
            fpy = lane_sum<TPA>(fpy);
            fpz = lane_sum<TPA>(fpz);
            if (sub != 0 || !have) continue;
            if (sp.write_f) {
                a.fx[p] = fpx;
                a.fy[p] = fpy;
                a.fz[p] = fpz;
            }
            if constexpr (MODE >= 1)
                integrate_atom<MODE>(a, t, B, sp, p, st, active, pi, ti, fpx, fpy, fpz);
        }
    }
}

namespace {

int device_attr(hipDeviceAttribute_t attr, int fallback) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fallback;
    if (hipDeviceGetAttribute(&v, attr, dev) != hipSuccess || v <= 0) return fallback;
    return v;
}

struct DevLimits {
    int cus = 0, lds_cu = 0, lds_block = 0;
};

const DevLimits& dev_limits() {
    static DevLimits lim[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    DevLimits& d = lim[dev];
    if (d.cus == 0) {
        d.lds_cu = device_attr(hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, 65536);
        d.lds_block = device_attr(hipDeviceAttributeMaxSharedMemoryPerBlock, 65536);
        d.cus = device_attr(hipDeviceAttributeMultiprocessorCount, 256);
    }
    return d;
}

// Persistent grid: min(requested, co-resident) workgroups per CU of kernel k (cached per kernel,
// block and LDS size) times the CU count, capped by the work.
unsigned persistent_grid(const void* k, int block, size_t lds, int wgs, int nchunks) {
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, size_t>, int> cache;
    int per_cu = 0;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find({k, block, lds});
        if (it != cache.end()) per_cu = it->second;
    }
    if (per_cu == 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k, block, lds) != hipSuccess || v <= 0)
            v = 1;
        per_cu = v;
        std::lock_guard<std::mutex> g(mu);
        cache[{k, block, lds}] = v;
    }
    if (wgs > 0 && wgs < per_cu) per_cu = wgs;
    const long long g = static_cast<long long>(per_cu) * dev_limits().cus;
    return static_cast<unsigned>(g < nchunks ? g : nchunks);
}

// LDS bytes one workgroup may take when wgs of them share a CU.
int lds_budget(int wgs) {
    const DevLimits& d = dev_limits();
    int b = d.lds_cu / (wgs > 0 ? wgs : 1);
    if (b > d.lds_block) b = d.lds_block;
    return b & ~255;
}

}  // namespace

LtWin lt_window(const TabDev& t, int wgs_d, int wgs_f) {
    LtWin w{};
    const int nk = t.ngroup > 1 ? t.nkind : 1;
    const int bf = lds_budget(wgs_f) / (16 * nk), bd = lds_budget(wgs_d) / (8 * nk);
    w.wf = t.ntab2 < bf ? t.ntab2 : bf;
    w.k0f = t.ntab2 - w.wf;
    w.wd = t.ntab2 < bd ? t.ntab2 : bd;
    w.k0d = t.ntab2 - w.wd;
    return w;
}

void launch_density_lt(int tpa, const LtShape& sh, const AtomsDev& a, const ListDev& l,
                       const TabDev& t, const BoxDev& box0, const BoxDev* boxes, int multibox,
                       int* err, hipStream_t s) {
    if (a.n <= 0) return;
    LtWin w = lt_window(t, sh.wgs_d, sh.wgs_f);
    w.sweep = sh.sweep;
    const bool mt = t.ngroup > 1;
    const size_t lds = static_cast<size_t>(w.wd) * (mt ? t.nkind : 1) * sizeof(double);
    const long long th = static_cast<long long>(a.n) * (tpa == 3 ? 2 : tpa);
    const int nchunks = static_cast<int>((th + sh.block_d - 1) / sh.block_d);
    const bool eam = t.pot_type != 0;
#define MSM_DL(MB, E, P)                                                                     \
    do {                                                                                     \
        auto k = mt ? k_density_lt<MB, 1, E, P> : k_density_lt<MB, 0, E, P>;                 \
        const unsigned g = persistent_grid(reinterpret_cast<const void*>(k), sh.block_d,     \
                                           lds, sh.wgs_d, nchunks);                          \
        hipLaunchKernelGGL(k, dim3(g), dim3(sh.block_d), lds, s, a, l, t, box0, boxes, w,    \
                           nchunks, err);                                                    \
    } while (0)
#define MSM_DLP(MB, E)                          \
    do {                                        \
        if (tpa == 1) MSM_DL(MB, E, 1);         \
        else if (tpa == 2) MSM_DL(MB, E, 2);    \
        else if (tpa == 3) MSM_DL(MB, E, 3);    \
        else MSM_DL(MB, E, 4);                  \
    } while (0)
    if (multibox) {
        if (eam) MSM_DLP(1, 1); else MSM_DLP(1, 0);
    } else {
        if (eam) MSM_DLP(0, 1); else MSM_DLP(0, 0);
    }
#undef MSM_DLP
#undef MSM_DL
}

void launch_force_lt(int tpa, const LtShape& sh, const AtomsDev& a, const ListDev& l,
                     const TabDev& t, const BoxDev& box0, const BoxDev* boxes, int multibox,
                     const StepParams& sp, hipStream_t s) {
    if (a.n <= 0) return;
    LtWin w = lt_window(t, sh.wgs_d, sh.wgs_f);
    w.sweep = sh.sweep;
    const bool mt = t.ngroup > 1;
    const size_t lds = static_cast<size_t>(w.wf) * (mt ? t.nkind : 1) * sizeof(double2);
    const long long th = static_cast<long long>(a.n) * (tpa == 3 ? 2 : tpa);
    const int nchunks = static_cast<int>((th + sh.block_f - 1) / sh.block_f);
#define MSM_FL(MB, M, P)                                                                     \
    do {                                                                                     \
        auto k = mt ? k_force_lt<MB, 1, M, P> : k_force_lt<MB, 0, M, P>;                     \
        const unsigned g = persistent_grid(reinterpret_cast<const void*>(k), sh.block_f,     \
                                           lds, sh.wgs_f, nchunks);                          \
        hipLaunchKernelGGL(k, dim3(g), dim3(sh.block_f), lds, s, a, l, t, box0, boxes, sp,   \
                           w, nchunks);                                                      \
    } while (0)
#define MSM_FLP(MB, M)                          \
    do {                                        \
        if (tpa == 1) MSM_FL(MB, M, 1);         \
        else if (tpa == 2) MSM_FL(MB, M, 2);    \
        else if (tpa == 3) MSM_FL(MB, M, 3);    \
        else MSM_FL(MB, M, 4);                  \
    } while (0)
#define MSM_FLM(MB)                             \
    do {                                        \
        if (sp.mode == 0) MSM_FLP(MB, 0);       \
        else if (sp.mode == 1) MSM_FLP(MB, 1);  \
        else MSM_FLP(MB, 2);                    \
    } while (0)
    if (multibox) MSM_FLM(1); else MSM_FLM(0);
#undef MSM_FLM
#undef MSM_FLP
#undef MSM_FL
}

}  // namespace msm

namespace msm {

namespace {

// Halo capacity and knot window of one brick pass sharing `budget` bytes: the capacity of the
// list build (scap), less if needed to leave a window of a third of the table.
void bt_fit(long budget, long bpa, long bpk, int ntab2, int scap, int& cap, int& wnd) {
    const long wmin = (ntab2 + 2) / 3;
    long c = std::min<long>(scap, (budget - (wmin + 2) * bpk) / bpa);
    if (c < 0) c = 0;
    const long wv = (budget - c * bpa) / bpk - 2;
    cap = static_cast<int>(c);
    wnd = static_cast<int>(std::max(0L, std::min<long>(ntab2, wv)));
}

}  // namespace

BtPlan bt_plan(const TabDev& t, const LtShape& sh, int scap) {
    BtPlan p{};
    const bool mt = t.ngroup > 1;
    const long nk = mt ? t.nkind : 1;
    const long fixed = static_cast<long>(sizeof(BrickLds)) + 1024;
    bt_fit(lds_budget(sh.wgs_d) - fixed, bt_atom_bytes(false, mt), 8 * nk, t.ntab2, scap,
           p.cap_d, p.w.wd);
    bt_fit(lds_budget(sh.wgs_f) - fixed, bt_atom_bytes(true, mt), 16 * nk, t.ntab2, scap,
           p.cap_f, p.w.wf);
    p.w.k0d = t.ntab2 - p.w.wd;
    p.w.k0f = t.ntab2 - p.w.wf;
    return p;
}

namespace {

size_t bt_lds(const BtPlan& p, const TabDev& t, bool force) {
    const bool mt = t.ngroup > 1;
    const int nk = mt ? t.nkind : 1;
    if (force)
        return static_cast<size_t>((nk * p.w.wf + 1) & ~1) * sizeof(double2) +
               static_cast<size_t>(p.cap_f) * bt_atom_bytes(true, mt);
    return static_cast<size_t>((nk * p.w.wd + 1) & ~1) * sizeof(double) +
           static_cast<size_t>(p.cap_d) * bt_atom_bytes(false, mt);
}

}  // namespace

void launch_density_bt(int tpa, const LtShape& sh, int nbricks, const int* brick_box,
                       const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev* boxes,
                       const int* start, int ncell_total, int* err, hipStream_t s) {
    if (a.n <= 0) return;
    launch_outside_atoms(a, start + ncell_total, 0, 0, 1, s);
    if (nbricks <= 0) return;
    const BtPlan pl = bt_plan(t, sh, l.scap);
    const size_t lds = bt_lds(pl, t, false);
    const bool mt = t.ngroup > 1, eam = t.pot_type != 0;
#define MSM_DB(MT, E, P)                                                                       \
    do {                                                                                       \
        auto k = k_density_bt<MT, E, P>;                                                       \
        const unsigned g = persistent_grid(reinterpret_cast<const void*>(k), sh.block_d, lds,  \
                                           sh.wgs_d, nbricks);                                 \
        hipLaunchKernelGGL(k, dim3(g), dim3(sh.block_d), lds, s, brick_box, nbricks, a, l, t,  \
                           boxes, start, pl.w, pl.cap_d, err);                                 \
    } while (0)
#define MSM_DBP(MT, E)                          \
    do {                                        \
        if (tpa == 1) MSM_DB(MT, E, 1);         \
        else if (tpa == 2) MSM_DB(MT, E, 2);    \
        else MSM_DB(MT, E, 4);                  \
    } while (0)
    if (mt) { if (eam) MSM_DBP(1, 1); else MSM_DBP(1, 0); }
    else    { if (eam) MSM_DBP(0, 1); else MSM_DBP(0, 0); }
#undef MSM_DBP
#undef MSM_DB
}

void launch_force_bt(int tpa, const LtShape& sh, int nbricks, const int* brick_box,
                     const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev* boxes,
                     const int* start, int ncell_total, const StepParams& sp, hipStream_t s) {
    if (a.n <= 0) return;
    launch_outside_atoms(a, start + ncell_total, sp.write_f, sp.mode == 2 ? 1 : 0, 0, s);
    if (nbricks <= 0) return;
    const BtPlan pl = bt_plan(t, sh, l.scap);
    const size_t lds = bt_lds(pl, t, true);
    const bool mt = t.ngroup > 1;
#define MSM_FB(MT, M, P)                                                                       \
    do {                                                                                       \
        auto k = k_force_bt<MT, M, P>;                                                         \
        const unsigned g = persistent_grid(reinterpret_cast<const void*>(k), sh.block_f, lds,  \
                                           sh.wgs_f, nbricks);                                 \
        hipLaunchKernelGGL(k, dim3(g), dim3(sh.block_f), lds, s, brick_box, nbricks, a, l, t,  \
                           boxes, start, sp, pl.w, pl.cap_f);                                  \
    } while (0)
#define MSM_FBP(MT, M)                          \
    do {                                        \
        if (tpa == 1) MSM_FB(MT, M, 1);         \
        else if (tpa == 2) MSM_FB(MT, M, 2);    \
        else MSM_FB(MT, M, 4);                  \
    } while (0)
#define MSM_FBM(MT)                             \
    do {                                        \
        if (sp.mode == 0) MSM_FBP(MT, 0);       \
        else if (sp.mode == 1) MSM_FBP(MT, 1);  \
        else MSM_FBP(MT, 2);                    \
    } while (0)
    if (mt) MSM_FBM(1); else MSM_FBM(0);
#undef MSM_FBM
#undef MSM_FBP
#undef MSM_FB
}

}  // namespace msm
