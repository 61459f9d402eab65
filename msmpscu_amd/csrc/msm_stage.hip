// msm_stage.hip -- density and force passes over LDS-staged bricks of cells, with the neighbour
// records of a brick staged once, coalesced, and every pair read from LDS.
//
// The global-list passes gather a 24/32-byte neighbour record from the L1/L2 for every pair; a
// wave's gather touches up to 64 distinct cache lines, so those passes are bound by the vector
// memory pipeline's line rate, not by HBM.  Here a workgroup takes one brick of at most 4^3 cells
// at a time (the bricks of the list build, whose per-pair halo slots `sidx` it reads), copies the
// brick's halo -- the brick plus one cell layer -- into LDS with row-contiguous loads, and then
// every home atom walks its list as 2-byte halo slots: one coalesced slot load per pair and LDS
// reads for the neighbour, so the per-pair global traffic is the slot stream alone.
//
// Periodic images are resolved once per staged atom instead of once per pair: a staged record
// holds the image of the atom nearest the centre of its halo cell, x - L*rint((x - c)/L) (one
// fma, exact for the image 0), which for every listed pair is the image the list build chose, so
// the pair separation is a plain difference.  For pairs whose image is 0 on both sides this is
// bit for bit the minimum-image difference of the other passes; across a periodic face the two
// roundings of (xi + kL) - (xj + k'L) replace the one of (xi - xj) - kL, a difference of an ulp
// of L in the separation.
//
// The persistent grid sweeps the bricks XCD by XCD: the workgroups of one XCD take interleaved
// bricks of that XCD's contiguous share, so the bricks in flight on an XCD are neighbours whose
// halos overlap in that XCD's L2.  Brick halos larger than the staging capacity (or bricks of a
// list without halo slots) take the global-list path with the minimum image per pair.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "msm_brick.h"

namespace msm {

namespace {

constexpr int kSbMaxBlock = 1024;
// slots loaded ahead per chunk of an atom's list (a lane issues them all before its first pair);
// the force pass holds more per pair, so its chunk is shorter (128 VGPRs at 1024 threads)
constexpr int kSbChunk = 32;
constexpr int kSbChunkF = 16;
// pairs per arithmetic group
constexpr int kSbGroupD = 4;
constexpr int kSbGroupF = 4;

// Brick order: XCD x (workgroup ids x, x+8, ...) owns the contiguous brick range
// [x*nb/8, (x+1)*nb/8) and its workgroups take that range's bricks round-robin, so the bricks in
// flight on one XCD are adjacent.  With fewer than 8 workgroups every brick is taken by plain
// striding.  Each brick is visited by exactly one workgroup.
__device__ __forceinline__ void sweep_range(int nb, int& first, int& step, int& end) {
    const int nblk = gridDim.x;
    if (nblk < 8) {
        first = blockIdx.x;
        step = nblk;
        end = nb;
        return;
    }
    const int x = blockIdx.x & 7, k = blockIdx.x >> 3;
    const int nx = (nblk - x + 7) >> 3;  // workgroups on XCD x
    first = static_cast<int>(static_cast<long long>(x) * nb / 8) + k;
    end = static_cast<int>(static_cast<long long>(x + 1) * nb / 8);
    step = nx;
}

// Image of coordinate v nearest the centre c of its halo cell (periodic directions; ilen = 0
// leaves free directions unchanged).
__device__ __forceinline__ double near_image(double v, double c, double len, double ilen) {
    return fma(-len, rint((v - c) * ilen), v);
}

// Centre of halo cell h (brick frame, unwrapped) in direction k.
__device__ __forceinline__ double halo_centre(const BoxDev& B, const BrickInfo& g, int hk, int k) {
    return B.lo[k] + (static_cast<double>(g.c0[k] - 1 + hk) + 0.5) * (B.len[k] / B.ncell[k]);
}

// 1/r and SK = sqrt(r) * CSI from r^2: hardware reciprocal square roots, each refined by one
// Newton step (relative error ~1e-16 after the step), sqrt(r) = rsqrt(1/r).
__device__ __forceinline__ double coord_fast(double r2, double csi, double& inv_r) {
    double y = __builtin_amdgcn_rsq(r2);
    y = y * fma(-0.5 * r2 * y, y, 1.5);
    inv_r = y;
    double z = __builtin_amdgcn_rsq(y);
    z = z * fma(-0.5 * y * z, z, 1.5);
    return z * csi;
}

}  // namespace

// Dynamic LDS: the POTB knot window (nk x wd doubles), then the staged halo as double4 records
// {x, y, z, -} (one LDS address per pair), then (MT) the types.
template <int MT, int EAM>
__global__ __launch_bounds__(kSbMaxBlock) void k_density_sb(const int* __restrict__ brick_box,
                                                            int nbricks, AtomsDev a, ListDev l,
                                                            TabDev t,
                                                            const BoxDev* __restrict__ boxes,
                                                            const int* __restrict__ start, LtWin w,
                                                            int cap, int* err) {
    extern __shared__ double s_dyn[];
    __shared__ BrickLds L;
    const int nk = MT ? t.nkind : 1;
    double* s_pb = s_dyn;
    const int nwin = (nk * w.wd + 3) & ~3;  // keep the records 32-byte aligned
    double4* lp = reinterpret_cast<double4*>(s_dyn + nwin);
    int* lty = reinterpret_cast<int*>(lp + cap);
    for (int k = threadIdx.x; k < nk * w.wd; k += blockDim.x) {
        const int kd = MT ? k / w.wd : t.kpair[0], q = MT ? k % w.wd : k;
        s_pb[k] = t.b2[static_cast<size_t>(kd) * t.ntab2 + w.k0d + q].x;
    }
    const int kmax = t.ntab2 - 2;
    int bi, bstep, bend;
    sweep_range(nbricks, bi, bstep, bend);
    for (; bi < bend; bi += bstep) {
        __syncthreads();  // the previous brick's halo is no longer read
        const BoxDev& B = boxes[brick_box[bi]];
        BrickInfo g;
        brick_info(B, bi, g);
        const int total = brick_halo(B, g, start, L.hcell, nullptr, L.hoff, L.homeoff, L.scan);
        const bool staged = total <= cap && l.scap > 0 && total <= l.scap;
        if (staged) {
            for_halo_slots(g, start, L.hcell, L.hoff, [&](int s, int q, int h) {
                // x, y, z only: w (F') of other atoms is being written by this pass
                const int h0 = h % g.H[0], h1 = (h / g.H[0]) % g.H[1], h2 = h / (g.H[0] * g.H[1]);
                const double2 xy = *reinterpret_cast<const double2*>(&a.pos[q]);
                lp[s] = make_double4(
                    near_image(xy.x, halo_centre(B, g, h0, 0), B.len[0], B.ilen[0]),
                    near_image(xy.y, halo_centre(B, g, h1, 1), B.len[1], B.ilen[1]),
                    near_image(a.pos[q].z, halo_centre(B, g, h2, 2), B.len[2], B.ilen[2]), 0.0);
                if constexpr (MT) lty[s] = a.ityp[q];
            });
        }
        __syncthreads();
        const int nhome = L.homeoff[g.nhc];
        for (int ha = threadIdx.x; ha - static_cast<int>(threadIdx.x) < nhome;
             ha += blockDim.x) {
            if (ha >= nhome) continue;
            int h, off;
            home_atom(ha, g, L.homeoff, h, off);
            const int p = start[L.hcell[h]] + off;
            const bool active = (a.statu[p] & STATU_ACTIVE) == STATU_ACTIVE;
            double den0 = 0.0;
            const int ti = MT ? a.ityp[p] : 0;
            if (active) {
                const int kv = l.kvois[p];
                if (staged) {
                    const int si = L.hoff[h] + off;
                    const double4 pi = lp[si];
                    const unsigned short* __restrict__ li = l.sidx + list_base(p, l.mx);
                    for (int c0 = 0; c0 < kv; c0 += kSbChunk) {
                        // every slot of the chunk is issued before the first pair (the list is
                        // allocated with a full chunk of slack past MXKVOIS; slots past kv are
                        // masked below)
                        int qs[kSbChunk];
#pragma unroll
                        for (int u = 0; u < kSbChunk; ++u)
                            qs[u] = li[static_cast<size_t>(c0 + u) * l.stride];
                        const int nc = min(kv - c0, kSbChunk);
#pragma unroll
                        for (int u0 = 0; u0 < kSbChunk; u0 += kSbGroupD) {
                            if (u0 >= nc) break;
                            unsigned fix = 0;
                            double sk_[kSbGroupD];
#pragma unroll
                            for (int u = u0; u < u0 + kSbGroupD; ++u) {
                                const bool use = u < nc;
                                const int q = use ? qs[u] : si;
                                const double4 pj = lp[q];
                                const double sx = pi.x - pj.x, sy = pi.y - pj.y, sz = pi.z - pj.z;
                                const double r2 = sx * sx + sy * sy + sz * sz;
                                const int row = MT ? t.kpair[ti * t.ngroup + lty[q]] * w.wd : 0;
                                double inv_r;
                                const double sk = coord_fast(use ? r2 : 1.0, t.csi, inv_r);
                                sk_[u - u0] = sk;
                                const int kk = min(static_cast<int>(sk), kmax);
                                const bool in = use && r2 <= t.ru2max;
                                const int kl = kk - w.k0d;
                                if (kl < 0 && in) fix |= 1u << (u - u0);
                                const int qq = row + max(kl, 0);
                                const double p0 = s_pb[qq], p1 = s_pb[qq + 1];
                                const double v = p0 + (sk - kk) * (p1 - p0);
                                den0 = den0 + ((in && kl >= 0) ? v : 0.0);
                            }
                            if (fix) {
                                // pairs below the LDS window: knots of the global table
#pragma unroll
                                for (int u = u0; u < u0 + kSbGroupD; ++u) {
                                    if (!(fix & (1u << (u - u0)))) continue;
                                    const int q = qs[u];
                                    const int kind =
                                        MT ? t.kpair[ti * t.ngroup + lty[q]] : t.kpair[0];
                                    const double sk = sk_[u - u0];
                                    const int kk = min(static_cast<int>(sk), kmax);
                                    const double2 e = t.b2[static_cast<size_t>(kind) * t.ntab2 + kk];
                                    den0 = den0 + (e.x + (sk - kk) * e.y);
                                }
                            }
                        }
                    }
                } else {
                    // halo too large to stage: neighbour records from global memory
                    const double4 pi = a.pos[p];
                    const int* __restrict__ li = l.indi + list_base(p, l.mx);
                    for (int iw = 0; iw < kv; ++iw) {
                        const int j = li[static_cast<size_t>(iw) * l.stride];
                        const double4 pj = a.pos[j];
                        const double sx = min_image(pi.x - pj.x, B.len[0], B.ilen[0]);
                        const double sy = min_image(pi.y - pj.y, B.len[1], B.ilen[1]);
                        const double sz = min_image(pi.z - pj.z, B.len[2], B.ilen[2]);
                        const double r2 = sx * sx + sy * sy + sz * sz;
                        if (!(r2 <= t.ru2max)) continue;
                        const int kind = MT ? t.kpair[ti * t.ngroup + a.ityp[j]] : t.kpair[0];
                        double inv_r;
                        const double sk = coord_fast(r2, t.csi, inv_r);
                        const int kk = min(static_cast<int>(sk), kmax);
                        const double2 e = t.b2[static_cast<size_t>(kind) * t.ntab2 + kk];
                        den0 = den0 + (e.x + (sk - kk) * e.y);
                    }
                }
                if constexpr (!EAM) {
                    if (den0 > 0.0) den0 = -0.5 / sqrt(den0);
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

// Dynamic LDS: the {FPOTR, FPOTB} knot window (nk x wf), then the staged halo as double4 records
// {x, y, z, F'} (cap of them), then (MT) the types.
template <int MT, int MODE>
__global__ __launch_bounds__(kSbMaxBlock) void k_force_sb(const int* __restrict__ brick_box,
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
    const int kmax = t.ntab2 - 2;
    int bi, bstep, bend;
    sweep_range(nbricks, bi, bstep, bend);
    for (; bi < bend; bi += bstep) {
        __syncthreads();
        const BoxDev& B = boxes[brick_box[bi]];
        BrickInfo g;
        brick_info(B, bi, g);
        const int total = brick_halo(B, g, start, L.hcell, nullptr, L.hoff, L.homeoff, L.scan);
        const bool staged = total <= cap && l.scap > 0 && total <= l.scap;
        if (staged) {
            for_halo_slots(g, start, L.hcell, L.hoff, [&](int s, int q, int h) {
                const int h0 = h % g.H[0], h1 = (h / g.H[0]) % g.H[1], h2 = h / (g.H[0] * g.H[1]);
                const double4 r = a.pos[q];
                lp[s] = make_double4(
                    near_image(r.x, halo_centre(B, g, h0, 0), B.len[0], B.ilen[0]),
                    near_image(r.y, halo_centre(B, g, h1, 1), B.len[1], B.ilen[1]),
                    near_image(r.z, halo_centre(B, g, h2, 2), B.len[2], B.ilen[2]), r.w);
                if constexpr (MT) lty[s] = a.ityp[q];
            });
        }
        __syncthreads();
        const int nhome = L.homeoff[g.nhc];
        for (int ha = threadIdx.x; ha - static_cast<int>(threadIdx.x) < nhome;
             ha += blockDim.x) {
            if (ha >= nhome) continue;
            int h, off;
            home_atom(ha, g, L.homeoff, h, off);
            const int p = start[L.hcell[h]] + off;
            const int st = a.statu[p];
            const bool active = (st & STATU_ACTIVE) == STATU_ACTIVE;
            const int ti = MT ? a.ityp[p] : 0;
            double fpx = 0.0, fpy = 0.0, fpz = 0.0;
            if (active) {
                const int kv = l.kvois[p];
                if (staged) {
                    const int si = L.hoff[h] + off;
                    const double4 pi = lp[si];
                    const double denki = pi.w;
                    const unsigned short* __restrict__ li = l.sidx + list_base(p, l.mx);
                    for (int c0 = 0; c0 < kv; c0 += kSbChunkF) {
                        int qs[kSbChunkF];
#pragma unroll
                        for (int u = 0; u < kSbChunkF; ++u)
                            qs[u] = li[static_cast<size_t>(c0 + u) * l.stride];
                        const int nc = min(kv - c0, kSbChunkF);
#pragma unroll
                        for (int u0 = 0; u0 < kSbChunkF; u0 += kSbGroupF) {
                            if (u0 >= nc) break;
#pragma unroll
                            for (int u = u0; u < u0 + kSbGroupF; ++u) {
                                const bool use = u < nc;
                                const int q = use ? qs[u] : si;
                                const double4 pj = lp[q];
                                const double sx = pi.x - pj.x, sy = pi.y - pj.y, sz = pi.z - pj.z;
                                const double r2 = sx * sx + sy * sy + sz * sz;
                                int row = 0, row1 = 0;
                                if constexpr (MT) {
                                    const int tj = lty[q];
                                    row = t.kpair[ti * t.ngroup + tj] * w.wf;
                                    row1 = t.kpair[tj * t.ngroup + ti] * w.wf;
                                }
                                double inv_r;
                                const double sk = coord_fast(use ? r2 : 1.0, t.csi, inv_r);
                                const int kk = min(static_cast<int>(sk), kmax);
                                const double dk = sk - kk;
                                const bool in = use && r2 <= t.ru2max;
                                const int kl = kk - w.k0f;
                                const int qq = max(kl, 0);
                                const double2 t0 = s_ft[row + qq], t1 = s_ft[row + qq + 1];
                                const double fb = t0.y + dk * (t1.y - t0.y);
                                double fb1 = fb;
                                if constexpr (MT) {
                                    const double b0 = s_ft[row1 + qq].y, b1 = s_ft[row1 + qq + 1].y;
                                    fb1 = b0 + dk * (b1 - b0);
                                }
                                // FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r
                                double ft = (t0.x + dk * (t1.x - t0.x)) * (inv_r * inv_r) +
                                            (fb * denki + fb1 * pj.w) * inv_r;
                                if (kl < 0 && in) {
                                    // below the LDS window: knots of the global table
                                    int kind = t.kpair[0], kind1 = t.kpair[0];
                                    if constexpr (MT) {
                                        const int tj = lty[q];
                                        kind = t.kpair[ti * t.ngroup + tj];
                                        kind1 = t.kpair[tj * t.ngroup + ti];
                                    }
                                    const double4 f = t.f4[static_cast<size_t>(kind) * t.ntab2 + kk];
                                    const double4 f1 =
                                        t.f4[static_cast<size_t>(kind1) * t.ntab2 + kk];
                                    ft = (f.x + dk * f.y) * (inv_r * inv_r) +
                                         ((f.z + dk * f.w) * denki + (f1.z + dk * f1.w) * pj.w) *
                                             inv_r;
                                }
                                const double fortot = in ? ft : 0.0;
                                fpx = fpx + fortot * sx;
                                fpy = fpy + fortot * sy;
                                fpz = fpz + fortot * sz;
                            }
                        }
                    }
                } else {
                    const double4 pg = a.pos[p];
                    const double denki = pg.w;
                    const int* __restrict__ li = l.indi + list_base(p, l.mx);
                    for (int iw = 0; iw < kv; ++iw) {
                        const int j = li[static_cast<size_t>(iw) * l.stride];
                        const double4 pj = a.pos[j];
                        const double sx = min_image(pg.x - pj.x, B.len[0], B.ilen[0]);
                        const double sy = min_image(pg.y - pj.y, B.len[1], B.ilen[1]);
                        const double sz = min_image(pg.z - pj.z, B.len[2], B.ilen[2]);
                        const double r2 = sx * sx + sy * sy + sz * sz;
                        if (!(r2 <= t.ru2max)) continue;
                        int kind = t.kpair[0], kind1 = t.kpair[0];
                        if constexpr (MT) {
                            const int tj = a.ityp[j];
                            kind = t.kpair[ti * t.ngroup + tj];
                            kind1 = t.kpair[tj * t.ngroup + ti];
                        }
                        double inv_r;
                        const double sk = coord_fast(r2, t.csi, inv_r);
                        const int kk = min(static_cast<int>(sk), kmax);
                        const double dk = sk - kk;
                        const double4 f = t.f4[static_cast<size_t>(kind) * t.ntab2 + kk];
                        const double4 f1 = t.f4[static_cast<size_t>(kind1) * t.ntab2 + kk];
                        const double fortot =
                            (f.x + dk * f.y) * (inv_r * inv_r) +
                            ((f.z + dk * f.w) * denki + (f1.z + dk * f1.w) * pj.w) * inv_r;
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
            if constexpr (MODE >= 1) {
                const double4 pg = a.pos[p];  // unshifted: the integrator's position
                integrate_atom<MODE>(a, t, B, sp, p, st, active, pg, ti, fpx, fpy, fpz);
            }
        }
    }
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

// Workgroups of kernel k that fit a CU at this block and LDS size (cached), capped by wgs.
unsigned sb_grid(const void* k, int block, size_t lds, int wgs, int nbricks) {
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
    const long long g = static_cast<long long>(per_cu) * sb_limits().cus;
    return static_cast<unsigned>(g < nbricks ? g : nbricks);
}

long sb_budget(int wgs) {
    const SbLimits& d = sb_limits();
    long b = d.lds_cu / (wgs > 0 ? wgs : 1);
    if (b > d.lds_block) b = d.lds_block;
    return (b & ~255L) - static_cast<long>(sizeof(BrickLds)) - 512;
}

}  // namespace

SbPlan sb_plan(const TabDev& t, const LtShape& sh, int scap) {
    SbPlan p{};
    const bool mt = t.ngroup > 1;
    const long nk = mt ? t.nkind : 1;
    const long ad = sizeof(double4) + (mt ? sizeof(int) : 0);
    const long af = sizeof(double4) + (mt ? sizeof(int) : 0);
    p.cap = scap;
    const long wd = (sb_budget(sh.wgs_d) - static_cast<long>(scap) * ad) / (8 * nk) - 4;
    const long wf = (sb_budget(sh.wgs_f) - static_cast<long>(scap) * af) / (16 * nk) - 2;
    p.w.wd = static_cast<int>(std::max(0L, std::min<long>(t.ntab2, wd)));
    p.w.wf = static_cast<int>(std::max(0L, std::min<long>(t.ntab2, wf)));
    p.w.k0d = t.ntab2 - p.w.wd;
    p.w.k0f = t.ntab2 - p.w.wf;
    // usable when both windows hold at least a quarter of the table (the pairs of a solid lie
    // in its top part; the rare closer pair reads the global table)
    p.ok = scap > 0 && 4 * p.w.wd >= t.ntab2 && 4 * p.w.wf >= t.ntab2;
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
    const size_t lds = static_cast<size_t>((nk * pl.w.wd + 3) & ~3) * sizeof(double) +
                       static_cast<size_t>(pl.cap) * (sizeof(double4) + (mt ? sizeof(int) : 0));
#define MSM_DSB(MT, E)                                                                         \
    do {                                                                                       \
        auto k = k_density_sb<MT, E>;                                                          \
        const unsigned g = sb_grid(reinterpret_cast<const void*>(k), sh.block_d, lds,          \
                                   sh.wgs_d, nbricks);                                         \
        hipLaunchKernelGGL(k, dim3(g), dim3(sh.block_d), lds, s, brick_box, nbricks, a, l, t,  \
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
    const size_t lds = static_cast<size_t>((nk * pl.w.wf + 1) & ~1) * sizeof(double2) +
                       static_cast<size_t>(pl.cap) * (sizeof(double4) + (mt ? sizeof(int) : 0));
#define MSM_FSB(MT, M)                                                                         \
    do {                                                                                       \
        auto k = k_force_sb<MT, M>;                                                            \
        const unsigned g = sb_grid(reinterpret_cast<const void*>(k), sh.block_f, lds,          \
                                   sh.wgs_f, nbricks);                                         \
        hipLaunchKernelGGL(k, dim3(g), dim3(sh.block_f), lds, s, brick_box, nbricks, a, l, t,  \
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
