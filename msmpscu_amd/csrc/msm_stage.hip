// msm_stage.hip -- density and force passes over LDS-staged bricks of cells: the neighbour
// records of a brick are staged once, with row-contiguous loads, and every pair reads LDS.
//
// The global-list passes gather a 32-byte neighbour record from the L1/L2 for every pair; a
// wave's gather touches up to 64 distinct cache lines, so those passes are bound by the vector
// memory pipeline's line rate, not by HBM.  Here one workgroup takes one brick of at most 4^3
// cells (the bricks of the list build, whose per-pair halo slots `sidx` it reads), copies the
// brick's halo -- the brick plus one cell layer -- into LDS, and then the home atoms walk their
// lists as 2-byte halo slots: one coalesced slot load per pair, the neighbour from LDS.
//
// Periodic images are resolved once per staged atom instead of once per pair: a staged record
// holds the image of the atom nearest the centre of its halo cell, x - L*rint((x - c)/L) (one
// fma, exact for the image 0), which for every listed pair is the image the list build chose.
// (A fixed per-cell shift would not do: the predictor wraps an atom that crosses a periodic face
// between rebuilds, and the atom keeps its cell slot until the next rebuild.)  The pair
// separation is then a plain difference.  For pairs whose image is 0 on both sides this is bit
// for bit the minimum-image difference of the other passes; across a periodic face the two
// roundings of (xi + kL) - (xj + k'L) replace the one of (xi - xj) - kL, an ulp of L apart.
//
// Lanes: a 512-thread workgroup serves 256 home atoms per round with two lanes each (lane s of an
// atom takes list slots s, s+2, ...).  The home atoms left after the full rounds take one more
// round with 2, 4, 8 or 16 lanes each, so that a brick a little larger than 256 atoms does not pay
// a second full-length round.  The lanes of an atom are adjacent; their sums meet by xor shuffles.
//
// The tables: the knots a solid's pairs use (the upper part of each kind's table) sit in LDS as a
// window; the rare closer pair reads the global table.  Density knots are the {POTB, dPOTB} pairs
// of the global table; force knots are {FPOTR, FPOTB}, with the slope taken from the next knot
// (the same subtraction the host makes for the global table).
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "msm_brick.h"

namespace msm {

namespace {

constexpr int kSbBlock = 512;        // threads per workgroup, one brick per workgroup
constexpr int kSbLanes = 2;          // lanes per home atom in the full rounds
constexpr int kSbGroup = 4;          // list slots per lane per arithmetic group

struct SbShared {
    BrickLds L;
    int hbase[kNlMaxHalo];           // sorted index of halo slot s of cell h: hbase[h] + s
    double hc[3][kNlMaxHalo];        // centre of halo cell h in the brick's unwrapped frame
    unsigned char homeh[64];         // halo cell of home cell c
    int kp[kMaxGroup * kMaxGroup];   // KPAIR (0-based kinds) for several atom types
};

// Brick of this workgroup, its halo table, the per-halo-cell atom bases and centres, and (when
// the halo fits cap) the halo cell of every slot.  Returns the number of halo atoms; all tables
// are valid on return.
__device__ __forceinline__ int sb_setup(const BoxDev& B, int bi, const int* __restrict__ start,
                                        const TabDev& t, int cap, BrickInfo& g, SbShared& S,
                                        unsigned char* smap) {
    brick_info(B, bi, g);
    const int total = brick_halo(B, g, start, S.L.hcell, nullptr, S.L.hoff, S.L.homeoff, S.L.scan);
    for (int k = threadIdx.x; k < g.nh; k += blockDim.x) {
        S.hbase[k] = S.L.hcell[k] >= 0 ? start[S.L.hcell[k]] - S.L.hoff[k] : 0;
        const int h3[3] = {k % g.H[0], (k / g.H[0]) % g.H[1], k / (g.H[0] * g.H[1])};
        for (int d = 0; d < 3; ++d)
            S.hc[d][k] = B.lo[d] + (static_cast<double>(g.c0[d] - 1 + h3[d]) + 0.5) *
                                       (B.len[d] / B.ncell[d]);
        if (total <= cap)
            for (int sl = S.L.hoff[k]; sl < S.L.hoff[k + 1]; ++sl)
                smap[sl] = static_cast<unsigned char>(k);
    }
    if (threadIdx.x < g.nhc) {
        const int c = threadIdx.x;
        const int ix = c % g.e[0], iy = (c / g.e[0]) % g.e[1], iz = c / (g.e[0] * g.e[1]);
        S.homeh[c] = static_cast<unsigned char>((ix + 1) + g.H[0] * ((iy + 1) + g.H[1] * (iz + 1)));
    }
    for (int k = threadIdx.x; k < t.ngroup * t.ngroup; k += blockDim.x) S.kp[k] = t.kpair[k];
    __syncthreads();  // the tables above are read by every wave
    return total;
}

// Image of coordinate v nearest the centre c of its halo cell (periodic directions; ilen = 0
// leaves free directions unchanged).
__device__ __forceinline__ double near_image(double v, double c, double len, double ilen) {
    return fma(-len, rint((v - c) * ilen), v);
}

// SK = sqrt(sqrt(r2)) * CSI (MD_FS_ForceTable.F90:248-252) and 1/r from refined reciprocal
// square roots (rsqrt_nr, within one ulp): 1/r = rsqrt(r2), sqrt(r) = rsqrt(1/r).
__device__ __forceinline__ double sb_coord(double r2, double csi, double& inv_r) {
    inv_r = rsqrt_nr(r2);
    return rsqrt_nr(inv_r) * csi;
}

// Stage the halo: record s is sorted atom hbase[h] + s of halo cell h, at its image nearest the
// cell's centre.  WITH_W: also the record's w (F'); without it only x, y, z are read, since the
// density pass writes w of other atoms meanwhile.  Thread tid takes slots tid, tid + blockDim,
// ... (consecutive slots are consecutive atoms of a cell row, so the loads coalesce), reads each
// slot's cell from the slot map, and issues all its loads before its stores.
template <int WITH_W, int MT>
__device__ __forceinline__ void sb_stage(const SbShared& S, const unsigned char* smap,
                                         const AtomsDev& a, const BoxDev& B, int total,
                                         double4* lp, int* lty) {
    constexpr int U = 3;
    for (int s0 = 0; s0 < total; s0 += U * static_cast<int>(blockDim.x)) {
        double4 r[U];
        int hh[U], ty[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int sl = s0 + u * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
            hh[u] = -1;
            if (sl >= total) continue;
            const int h = smap[sl];
            hh[u] = h;
            const int q = S.hbase[h] + sl;
            if constexpr (WITH_W) {
                r[u] = a.pos[q];
            } else {
                const double2 xy = *reinterpret_cast<const double2*>(&a.pos[q]);
                r[u] = make_double4(xy.x, xy.y, a.pos[q].z, 0.0);
            }
            if constexpr (MT) ty[u] = a.ityp[q];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (hh[u] < 0) continue;
            const int sl = s0 + u * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
            const int h = hh[u];
            lp[sl] = make_double4(near_image(r[u].x, S.hc[0][h], B.len[0], B.ilen[0]),
                                  near_image(r[u].y, S.hc[1][h], B.len[1], B.ilen[1]),
                                  near_image(r[u].z, S.hc[2][h], B.len[2], B.ilen[2]), r[u].w);
            if constexpr (MT) lty[sl] = ty[u];
        }
    }
}

// Home atom number u of the brick -> halo cell h and offset within that cell.
__device__ __forceinline__ void sb_home(int u, const BrickInfo& g, const SbShared& S, int& h,
                                        int& off) {
    int lo = 0, hi = g.nhc;  // homeoff[lo] <= u < homeoff[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (S.L.homeoff[mid] <= u) lo = mid; else hi = mid;
    }
    h = S.homeh[lo];
    off = u - S.L.homeoff[lo];
}

template <int L>
__device__ __forceinline__ double sb_lane_sum(double v) {
#pragma unroll
    for (int m = 1; m < L; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// rho_i over the list slots sub, sub + L, ... of one home atom (staged halo).  Per group of
// kSbGroup slots: the slot loads of the next group are issued first, then every neighbour record,
// then the arithmetic without branches; pairs below the LDS window are added from the global
// table after the group (rare in a solid).
template <int MT, int L>
__device__ __forceinline__ double sb_density_sum(const double4* lp, const int* lty,
                                                 const double2* win, const SbShared& S,
                                                 const TabDev& t, const LtWin& w,
                                                 const unsigned short* __restrict__ li,
                                                 int stride, int kv, int sub, int si,
                                                 const double4& pi, int ti, int kmax) {
    double den = 0.0;
    unsigned short nxt[kSbGroup];
#pragma unroll
    for (int u = 0; u < kSbGroup; ++u) {
        const int k = sub + u * L;
        nxt[u] = k < kv ? li[static_cast<size_t>(k) * stride] : 0;
    }
    for (int k0 = sub; k0 < kv; k0 += kSbGroup * L) {
        int q[kSbGroup];
#pragma unroll
        for (int u = 0; u < kSbGroup; ++u) q[u] = k0 + u * L < kv ? nxt[u] : si;
#pragma unroll
        for (int u = 0; u < kSbGroup; ++u) {
            const int k = k0 + (kSbGroup + u) * L;
            nxt[u] = k < kv ? li[static_cast<size_t>(k) * stride] : 0;
        }
        double4 pj[kSbGroup];
        int row[kSbGroup];
#pragma unroll
        for (int u = 0; u < kSbGroup; ++u) {
            pj[u] = lp[q[u]];
            row[u] = MT ? S.kp[ti * t.ngroup + lty[q[u]]] : 0;
        }
        double sk[kSbGroup];
        int kk[kSbGroup];
        unsigned fix = 0;
#pragma unroll
        for (int u = 0; u < kSbGroup; ++u) {
            const double sx = pi.x - pj[u].x, sy = pi.y - pj[u].y, sz = pi.z - pj[u].z;
            const double r2 = sx * sx + sy * sy + sz * sz;
            const bool in = k0 + u * L < kv && r2 <= t.ru2max;
            double inv_r;
            sk[u] = sb_coord(in ? r2 : 1.0, t.csi, inv_r);
            kk[u] = min(static_cast<int>(sk[u]), kmax);
            const int kl = kk[u] - w.k0d;
            const double2 e = win[row[u] * w.wd + max(kl, 0)];
            const double v = e.x + (sk[u] - kk[u]) * e.y;
            den = den + ((in && kl >= 0) ? v : 0.0);
            if (in && kl < 0) fix |= 1u << u;
        }
        if (fix) {
#pragma unroll
            for (int u = 0; u < kSbGroup; ++u) {
                if (!(fix & (1u << u))) continue;
                const int gk = MT ? row[u] : t.kpair[0];
                const double2 eg = t.b2[static_cast<size_t>(gk) * t.ntab2 + kk[u]];
                den = den + (eg.x + (sk[u] - kk[u]) * eg.y);
            }
        }
    }
    return den;
}

// One round of home atoms [base, base + nat) with L lanes each (density).
template <int MT, int EAM, int L>
__device__ __forceinline__ void sb_density_round(int base, int nat, bool staged,
                                                 const BrickInfo& g, const SbShared& S,
                                                 const double4* lp, const int* lty,
                                                 const double2* win, AtomsDev& a,
                                                 const ListDev& l, const TabDev& t,
                                                 const LtWin& w, const BoxDev& B, int kmax,
                                                 int* err) {
    const int idx = threadIdx.x / L, sub = threadIdx.x % L;
    if (idx >= nat) return;
    int h, off;
    sb_home(base + idx, g, S, h, off);
    const int si = S.L.hoff[h] + off;
    const int p = S.hbase[h] + si;
    const int ti = MT ? a.ityp[p] : 0;
    const bool active = (a.statu[p] & STATU_ACTIVE) == STATU_ACTIVE;
    double den = 0.0;
    if (active) {
        const int kv = l.kvois[p];
        if (staged) {
            den = sb_density_sum<MT, L>(lp, lty, win, S, t, w, l.sidx + list_base(p, l.mx),
                                        l.stride, kv, sub, si, lp[si], ti, kmax);
        } else {
            // halo too large to stage: neighbour records from global memory
            const double4 pi = a.pos[p];
            const int* __restrict__ li = l.indi + list_base(p, l.mx);
            for (int k = sub; k < kv; k += L) {
                const int j = li[static_cast<size_t>(k) * l.stride];
                const double4 pj = a.pos[j];
                const double sx = min_image(pi.x - pj.x, B.len[0], B.ilen[0]);
                const double sy = min_image(pi.y - pj.y, B.len[1], B.ilen[1]);
                const double sz = min_image(pi.z - pj.z, B.len[2], B.ilen[2]);
                const double r2 = sx * sx + sy * sy + sz * sz;
                if (!(r2 <= t.ru2max)) continue;
                const int kind = MT ? t.kpair[ti * t.ngroup + a.ityp[j]] : t.kpair[0];
                double inv_r;
                const double sk = table_coord(r2, t.csi, inv_r);
                const int kk = min(static_cast<int>(sk), kmax);
                const double2 e = t.b2[static_cast<size_t>(kind) * t.ntab2 + kk];
                den = den + (e.x + (sk - kk) * e.y);
            }
        }
    }
    den = sb_lane_sum<L>(den);
    if (sub != 0) return;
    if (active) {
        if constexpr (!EAM) {
            if (den > 0.0) den = fs_fprime(den);
        } else {
            if (den > 0.0) {
                const double sk = den / t.rhod + 1.0;
                int kk = static_cast<int>(sk + 0.000001);
                if (kk > t.nembd) {
                    atomicOr(&err[kErrEmbd], 1);
                    kk = t.nembd;
                }
                const double4 m = t.m4[t.kembd[ti] * t.nembd2 + kk];
                den = m.z + (sk - kk) * m.w;
            }
        }
    }
    a.pos[p].w = den;
}

// F_i over the list slots sub, sub + L, ... of one home atom (staged halo, w = F'); grouped as
// the density sum.
template <int MT, int L>
__device__ __forceinline__ void sb_force_sum(const double4* lp, const int* lty,
                                             const double2* win, const SbShared& S,
                                             const TabDev& t, const LtWin& w,
                                             const unsigned short* __restrict__ li, int stride,
                                             int kv, int sub, int si, const double4& pi, int ti,
                                             int kmax, double& fpx, double& fpy, double& fpz) {
    unsigned short nxt[kSbGroup];
#pragma unroll
    for (int u = 0; u < kSbGroup; ++u) {
        const int k = sub + u * L;
        nxt[u] = k < kv ? li[static_cast<size_t>(k) * stride] : 0;
    }
    for (int k0 = sub; k0 < kv; k0 += kSbGroup * L) {
        int q[kSbGroup];
#pragma unroll
        for (int u = 0; u < kSbGroup; ++u) q[u] = k0 + u * L < kv ? nxt[u] : si;
#pragma unroll
        for (int u = 0; u < kSbGroup; ++u) {
            const int k = k0 + (kSbGroup + u) * L;
            nxt[u] = k < kv ? li[static_cast<size_t>(k) * stride] : 0;
        }
        double4 pj[kSbGroup];
        int kind[kSbGroup], kind1[kSbGroup];
#pragma unroll
        for (int u = 0; u < kSbGroup; ++u) {
            pj[u] = lp[q[u]];
            kind[u] = 0;
            kind1[u] = 0;
            if constexpr (MT) {
                const int tj = lty[q[u]];
                kind[u] = S.kp[ti * t.ngroup + tj];
                kind1[u] = S.kp[tj * t.ngroup + ti];
            }
        }
        unsigned fix = 0;
#pragma unroll
        for (int u = 0; u < kSbGroup; ++u) {
            const double sx = pi.x - pj[u].x, sy = pi.y - pj[u].y, sz = pi.z - pj[u].z;
            const double r2 = sx * sx + sy * sy + sz * sz;
            const bool in = k0 + u * L < kv && r2 <= t.ru2max;
            double inv_r;
            const double sk = sb_coord(in ? r2 : 1.0, t.csi, inv_r);
            const int kk = min(static_cast<int>(sk), kmax);
            const double dk = sk - kk;
            const int kl = kk - w.k0f;
            const int qq = max(kl, 0);
            const double2 t0 = win[kind[u] * w.wf + qq], t1 = win[kind[u] * w.wf + qq + 1];
            const double fb = t0.y + dk * (t1.y - t0.y);
            double fb1 = fb;
            if constexpr (MT) {
                const double b0 = win[kind1[u] * w.wf + qq].y;
                const double b1 = win[kind1[u] * w.wf + qq + 1].y;
                fb1 = b0 + dk * (b1 - b0);
            }
            // FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r
            const double ft = (t0.x + dk * (t1.x - t0.x)) * (inv_r * inv_r) +
                              (fb * pi.w + fb1 * pj[u].w) * inv_r;
            const double fortot = (in && kl >= 0) ? ft : 0.0;
            fpx = fpx + fortot * sx;
            fpy = fpy + fortot * sy;
            fpz = fpz + fortot * sz;
            if (in && kl < 0) fix |= 1u << u;
        }
        if (fix) {
            // below the LDS window: the global table
#pragma unroll
            for (int u = 0; u < kSbGroup; ++u) {
                if (!(fix & (1u << u))) continue;
                const double sx = pi.x - pj[u].x, sy = pi.y - pj[u].y, sz = pi.z - pj[u].z;
                const double r2 = sx * sx + sy * sy + sz * sz;
                double inv_r;
                const double sk = sb_coord(r2, t.csi, inv_r);
                const int kk = min(static_cast<int>(sk), kmax);
                const double dk = sk - kk;
                const int gk = MT ? kind[u] : t.kpair[0], gk1 = MT ? kind1[u] : t.kpair[0];
                const double4 f = t.f4[static_cast<size_t>(gk) * t.ntab2 + kk];
                const double4 f1 = t.f4[static_cast<size_t>(gk1) * t.ntab2 + kk];
                const double fortot =
                    (f.x + dk * f.y) * (inv_r * inv_r) +
                    ((f.z + dk * f.w) * pi.w + (f1.z + dk * f1.w) * pj[u].w) * inv_r;
                fpx = fpx + fortot * sx;
                fpy = fpy + fortot * sy;
                fpz = fpz + fortot * sz;
            }
        }
    }
}

// One round of home atoms [base, base + nat) with L lanes each (force, then the fused
// integrator of MODE on the atom's first lane).
template <int MT, int MODE, int L>
__device__ __forceinline__ void sb_force_round(int base, int nat, bool staged, const BrickInfo& g,
                                               const SbShared& S, const double4* lp,
                                               const int* lty, const double2* win, AtomsDev& a,
                                               const ListDev& l, const TabDev& t, const LtWin& w,
                                               const BoxDev& B, const StepParams& sp, int kmax) {
    const int idx = threadIdx.x / L, sub = threadIdx.x % L;
    if (idx >= nat) return;
    int h, off;
    sb_home(base + idx, g, S, h, off);
    const int si = S.L.hoff[h] + off;
    const int p = S.hbase[h] + si;
    const int st = a.statu[p];
    const bool active = (st & STATU_ACTIVE) == STATU_ACTIVE;
    const int ti = MT ? a.ityp[p] : 0;
    double fpx = 0.0, fpy = 0.0, fpz = 0.0;
    if (active) {
        const int kv = l.kvois[p];
        if (staged) {
            sb_force_sum<MT, L>(lp, lty, win, S, t, w, l.sidx + list_base(p, l.mx), l.stride,
                                kv, sub, si, lp[si], ti, kmax, fpx, fpy, fpz);
        } else {
            const double4 pi = a.pos[p];
            const int* __restrict__ li = l.indi + list_base(p, l.mx);
            for (int k = sub; k < kv; k += L) {
                const int j = li[static_cast<size_t>(k) * l.stride];
                const double4 pj = a.pos[j];
                const double sx = min_image(pi.x - pj.x, B.len[0], B.ilen[0]);
                const double sy = min_image(pi.y - pj.y, B.len[1], B.ilen[1]);
                const double sz = min_image(pi.z - pj.z, B.len[2], B.ilen[2]);
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
                const int kk = min(static_cast<int>(sk), kmax);
                const double dk = sk - kk;
                const double4 f = t.f4[static_cast<size_t>(kind) * t.ntab2 + kk];
                const double4 f1 = t.f4[static_cast<size_t>(kind1) * t.ntab2 + kk];
                const double fortot =
                    (f.x + dk * f.y) * (inv_r * inv_r) +
                    ((f.z + dk * f.w) * pi.w + (f1.z + dk * f1.w) * pj.w) * inv_r;
                fpx = fpx + fortot * sx;
                fpy = fpy + fortot * sy;
                fpz = fpz + fortot * sz;
            }
        }
    }
    fpx = sb_lane_sum<L>(fpx);
    fpy = sb_lane_sum<L>(fpy);
    fpz = sb_lane_sum<L>(fpz);
    if (sub != 0) return;
    if (sp.write_f) {
        a.fx[p] = fpx;
        a.fy[p] = fpy;
        a.fz[p] = fpz;
    }
    if constexpr (MODE >= 1) {
        const double4 pg = a.pos[p];  // unshifted: the integrator's position
        integrate_atom<MODE>(a, t, B, sp, p, st, active, pg, ti, fpx, fpy, fpz);
    }
}

// Home atoms of a brick in rounds: full rounds of kSbBlock / kSbLanes atoms, then the rest with
// the most lanes per atom (up to 16) that one round can give them.
template <typename F2, typename F4, typename F8, typename F16>
__device__ __forceinline__ void sb_rounds(int nhome, F2&& r2, F4&& r4, F8&& r8, F16&& r16) {
    constexpr int full = kSbBlock / kSbLanes;
    int base = 0;
    for (; nhome - base > full / 2; base += full) r2(base, min(full, nhome - base));
    const int m = nhome - base;
    if (m <= 0) return;
    if (m <= kSbBlock / 16) r16(base, m);
    else if (m <= kSbBlock / 8) r8(base, m);
    else r4(base, m);
}

}  // namespace

// Dynamic LDS: the density window (nk x wd {POTB, dPOTB} knots), the staged halo as double4
// records {x, y, z, -} (cap of them), then (MT) the types, then the slot -> halo cell map.
template <int MT, int EAM>
__global__ __launch_bounds__(kSbBlock, 4) void k_density_sb(const int* __restrict__ brick_box,
                                                         int nbricks, AtomsDev a, ListDev l,
                                                         TabDev t,
                                                         const BoxDev* __restrict__ boxes,
                                                         const int* __restrict__ start, LtWin w,
                                                         int cap, int* err) {
    extern __shared__ double4 s_dyn4[];
    __shared__ SbShared S;
    const int nk = MT ? t.nkind : 1;
    double2* win = reinterpret_cast<double2*>(s_dyn4);
    double4* lp = s_dyn4 + (nk * w.wd + 1) / 2;
    int* lty = reinterpret_cast<int*>(lp + cap);
    unsigned char* smap = reinterpret_cast<unsigned char*>(lty + (MT ? cap : 0));
    for (int k = threadIdx.x; k < nk * w.wd; k += blockDim.x) {
        const int kd = MT ? k / w.wd : t.kpair[0], q = MT ? k % w.wd : k;
        win[k] = t.b2[static_cast<size_t>(kd) * t.ntab2 + w.k0d + q];
    }
    const int kmax = t.ntab2 - 2;
    const int bi = xcd_block(gridDim.x);
    if (bi >= nbricks) return;
    const BoxDev& B = boxes[brick_box[bi]];
    BrickInfo g;
    const int total = sb_setup(B, bi, start, t, cap, g, S, smap);
    const bool staged = total <= cap && l.scap > 0 && total <= l.scap;
    if (staged) sb_stage<0, MT>(S, smap, a, B, total, lp, lty);
    __syncthreads();
    auto round = [&](auto lanes, int base, int nat) {
        sb_density_round<MT, EAM, decltype(lanes)::value>(base, nat, staged, g, S, lp, lty, win,
                                                          a, l, t, w, B, kmax, err);
    };
    sb_rounds(
        S.L.homeoff[g.nhc],
        [&](int b, int m) { round(std::integral_constant<int, 2>{}, b, m); },
        [&](int b, int m) { round(std::integral_constant<int, 4>{}, b, m); },
        [&](int b, int m) { round(std::integral_constant<int, 8>{}, b, m); },
        [&](int b, int m) { round(std::integral_constant<int, 16>{}, b, m); });
}

// Dynamic LDS: the force window (nk x wf {FPOTR, FPOTB} knots), the staged halo as double4
// records {x, y, z, F'} (cap of them), then (MT) the types, then the slot -> halo cell map.
template <int MT, int MODE>
__global__ __launch_bounds__(kSbBlock, 4) void k_force_sb(const int* __restrict__ brick_box,
                                                       int nbricks, AtomsDev a, ListDev l,
                                                       TabDev t, const BoxDev* __restrict__ boxes,
                                                       const int* __restrict__ start,
                                                       StepParams sp, LtWin w, int cap) {
    extern __shared__ double4 s_dyn4[];
    __shared__ SbShared S;
    const int nk = MT ? t.nkind : 1;
    double2* win = reinterpret_cast<double2*>(s_dyn4);
    double4* lp = s_dyn4 + (nk * w.wf + 1) / 2;
    int* lty = reinterpret_cast<int*>(lp + cap);
    unsigned char* smap = reinterpret_cast<unsigned char*>(lty + (MT ? cap : 0));
    for (int k = threadIdx.x; k < nk * w.wf; k += blockDim.x) {
        const int kd = MT ? k / w.wf : t.kpair[0], q = MT ? k % w.wf : k;
        const double4 f = t.f4[static_cast<size_t>(kd) * t.ntab2 + w.k0f + q];
        win[k] = make_double2(f.x, f.z);
    }
    const int kmax = t.ntab2 - 2;
    const int bi = xcd_block(gridDim.x);
    if (bi >= nbricks) return;
    const BoxDev& B = boxes[brick_box[bi]];
    BrickInfo g;
    const int total = sb_setup(B, bi, start, t, cap, g, S, smap);
    const bool staged = total <= cap && l.scap > 0 && total <= l.scap;
    if (staged) sb_stage<1, MT>(S, smap, a, B, total, lp, lty);
    __syncthreads();
    auto round = [&](auto lanes, int base, int nat) {
        sb_force_round<MT, MODE, decltype(lanes)::value>(base, nat, staged, g, S, lp, lty, win, a,
                                                         l, t, w, B, sp, kmax);
    };
    sb_rounds(
        S.L.homeoff[g.nhc],
        [&](int b, int m) { round(std::integral_constant<int, 2>{}, b, m); },
        [&](int b, int m) { round(std::integral_constant<int, 4>{}, b, m); },
        [&](int b, int m) { round(std::integral_constant<int, 8>{}, b, m); },
        [&](int b, int m) { round(std::integral_constant<int, 16>{}, b, m); });
}

namespace {

int sb_device_attr(hipDeviceAttribute_t attr, int fallback) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fallback;
    if (hipDeviceGetAttribute(&v, attr, dev) != hipSuccess || v <= 0) return fallback;
    return v;
}

struct SbLimits {
    int cus = 0, lds_cu = 0, lds_block = 0;
};

const SbLimits& sb_limits() {
    static SbLimits lim[64];
    static std::mutex mu;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    std::lock_guard<std::mutex> g(mu);
    SbLimits& d = lim[dev];
    if (d.cus == 0) {
        d.lds_cu = sb_device_attr(hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, 65536);
        d.lds_block = sb_device_attr(hipDeviceAttributeMaxSharedMemoryPerBlock, 65536);
        d.cus = sb_device_attr(hipDeviceAttributeMultiprocessorCount, 256);
    }
    return d;
}

// LDS bytes of one workgroup when wgs of them share a CU, less the static tables.
long sb_budget(int wgs) {
    const SbLimits& d = sb_limits();
    long b = d.lds_cu / (wgs > 0 ? wgs : 1);
    if (b > d.lds_block) b = d.lds_block;
    return (b & ~255L) - static_cast<long>(sizeof(SbShared)) - 512;
}

// Knot window of `per` bytes per knot and kind in `bytes`: the top of each kind's table.
int sb_window(long bytes, long per, long nk, int ntab2) {
    const long w = bytes / (per * nk) - 2;
    return static_cast<int>(std::max(0L, std::min<long>(ntab2, w)));
}

}  // namespace

SbPlan sb_plan(const TabDev& t, const LtShape& sh, int scap) {
    (void)sh;
    SbPlan p{};
    const bool mt = t.ngroup > 1;
    const long nk = mt ? t.nkind : 1;
    const long rec = static_cast<long>(scap) * (sizeof(double4) + (mt ? sizeof(int) : 0) + 1);
    p.cap = scap;
    // two workgroups per CU when both windows keep a quarter of the table beside the halo, else one
    for (int wgs : {2, 1}) {
        const long room = sb_budget(wgs) - rec;
        p.w.wd = sb_window(room, sizeof(double2), nk, t.ntab2);
        p.w.wf = sb_window(room, sizeof(double2), nk, t.ntab2);
        p.w.k0d = t.ntab2 - p.w.wd;
        p.w.k0f = t.ntab2 - p.w.wf;
        p.ok = scap > 0 && 4 * p.w.wd >= t.ntab2 && 4 * p.w.wf >= t.ntab2;
        if (p.ok) break;
    }
    return p;
}

void launch_density_sb(const LtShape& sh, int nbricks, const int* brick_box, const AtomsDev& a,
                       const ListDev& l, const TabDev& t, const BoxDev* boxes, const int* start,
                       int ncell_total, int* err, hipStream_t s) {
    if (a.n <= 0) return;
    launch_outside_atoms(a, start + ncell_total, 0, 0, 1, s);
    if (nbricks <= 0) return;
    const SbPlan pl = sb_plan(t, sh, l.scap);
    const bool mt = t.ngroup > 1, eam = t.pot_type != 0;
    const int nk = mt ? t.nkind : 1;
    const size_t lds = static_cast<size_t>((nk * pl.w.wd + 1) / 2) * sizeof(double4) +
                       static_cast<size_t>(pl.cap) * (sizeof(double4) + (mt ? sizeof(int) : 0) + 1);
#define MSM_DSB(MT, E)                                                                         \
    do {                                                                                       \
        auto k = k_density_sb<MT, E>;                                                          \
        hipLaunchKernelGGL(k, dim3(nbricks), dim3(kSbBlock), lds, s, brick_box, nbricks, a, l, t, \
                           boxes, start, pl.w, pl.cap, err);                                   \
    } while (0)
    if (mt) { if (eam) MSM_DSB(1, 1); else MSM_DSB(1, 0); }
    else    { if (eam) MSM_DSB(0, 1); else MSM_DSB(0, 0); }
#undef MSM_DSB
}

void launch_force_sb(const LtShape& sh, int nbricks, const int* brick_box, const AtomsDev& a,
                     const ListDev& l, const TabDev& t, const BoxDev* boxes, const int* start,
                     int ncell_total, const StepParams& sp, hipStream_t s) {
    if (a.n <= 0) return;
    launch_outside_atoms(a, start + ncell_total, sp.write_f, sp.mode == 2 ? 1 : 0, 0, s);
    if (nbricks <= 0) return;
    const SbPlan pl = sb_plan(t, sh, l.scap);
    const bool mt = t.ngroup > 1;
    const int nk = mt ? t.nkind : 1;
    const size_t lds = static_cast<size_t>((nk * pl.w.wf + 1) / 2) * sizeof(double4) +
                       static_cast<size_t>(pl.cap) * (sizeof(double4) + (mt ? sizeof(int) : 0) + 1);
#define MSM_FSB(MT, M)                                                                         \
    do {                                                                                       \
        auto k = k_force_sb<MT, M>;                                                            \
        hipLaunchKernelGGL(k, dim3(nbricks), dim3(kSbBlock), lds, s, brick_box, nbricks, a, l, t, \
                           boxes, start, sp, pl.w, pl.cap);                                    \
    } while (0)
#define MSM_FSBM(MT)                            \
    do {                                        \
        if (sp.mode == 0) MSM_FSB(MT, 0);       \
        else if (sp.mode == 1) MSM_FSB(MT, 1);  \
        else MSM_FSB(MT, 2);                    \
    } while (0)
    if (mt) MSM_FSBM(1); else MSM_FSBM(0);
#undef MSM_FSBM
#undef MSM_FSB
}

}  // namespace msm
