// msm_api.hip -- C-ABI entry points of libmsm_hip.so (declared in include/msm_hip.h): device
// context, table / box / atom upload, neighbour-list build, force, energy, integrator and the
// graph-replayed step loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "msm_hip.h"
#include "msm_internal.h"

using msm::AtomsDev;
using msm::BoxDev;
using msm::ListDev;
using msm::SortBufs;
using msm::StepParams;
using msm::TabDev;

namespace {

struct AtomSet {  // one set of per-atom arrays (cell-sorted order)
    double* v = nullptr;    // 3 * n_alloc
    double* f = nullptr;    // 3 * n_alloc
    double* dis = nullptr;  // 3 * n_alloc
    int* ityp = nullptr;
    int* statu = nullptr;
    int* boxid = nullptr;
    int* gid = nullptr;
    int* cell = nullptr;
};

// A captured segment bakes in the step parameters (h enters the force kernels by value), so the
// time step's bit pattern is part of the key: a run with another h captures its own graphs.
struct GraphKey {
    int set, pos, rebuild, nsteps, last, track;
    uint64_t hbits;
    bool operator<(const GraphKey& o) const {
        return std::tie(set, pos, rebuild, nsteps, last, track, hbits) <
               std::tie(o.set, o.pos, o.rebuild, o.nsteps, o.last, o.track, o.hbits);
    }
};

}  // namespace

struct msm_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // force table
    bool has_table = false;
    TabDev tab{};
    double2* d_b2 = nullptr;
    double4 *d_f4 = nullptr, *d_e4 = nullptr, *d_m4 = nullptr;
    size_t cap_b2 = 0, cap_f4 = 0, cap_e4 = 0, cap_m4 = 0;  // elements allocated
    int tab_ngroup = 0;

    // control
    bool has_ctrl = false;
    int ngroup = 0, mx = 0;
    std::vector<double> ru, nb_rm, cm;
    double* d_rnl2 = nullptr;

    // boxes
    bool has_boxes = false;
    std::vector<BoxDev> boxes;
    BoxDev* d_boxes = nullptr;
    int ncell_total = 0;

    // atoms
    int n = 0, n_alloc = 0;
    double4* pos[2] = {nullptr, nullptr};
    int pcur = 0;  // index of the current position buffer
    AtomSet set[2];
    int scur = 0;  // index of the current atom set
    double* epot = nullptr;
    bool has_atoms = false;

    // sort + list
    SortBufs sb{};
    int sort_cap = 0;
    int* kvois = nullptr;
    int* indi = nullptr;
    unsigned short* sidx = nullptr;
    int* cidx = nullptr;  // in-cutoff pairs of the current step (density -> force pass)
    int* ckv = nullptr;
    unsigned long long* cmask = nullptr;  // in-cutoff slot masks per lane (LDS-table passes)
    size_t indi_cap = 0;
    bool has_list = false;

    // flags and scratch
    int* d_err = nullptr;
    int* h_err = nullptr;
    double* d_ebox = nullptr;
    unsigned long long* d_stats = nullptr;
    double* d_boxwork = nullptr;  // per-box scratch (4 doubles a box)
    size_t boxwork_cap = 0;
    unsigned long long draws = 0;  // msm_thermalize calls so far (the generator's draw number)

    // options
    int track_dis = 1;
    int timing = 0;
    int nlist_kind = 0;    // list build: 0 LDS bricks sized by occupancy, 1 per-atom global
                           // scan, 2 small LDS bricks, 3 wide LDS bricks
    int force_path = 2;    // 0: LDS bricks; 1, 2, 4, 8: global list with that many lanes per atom;
                           // 3: two lanes per atom with lane-pair gathers
    int table_lds = 1;     // table knots in LDS (global-list paths 1, 2, 4; brick path 0)
    int bt_lanes = 1;      // lanes per home atom of the brick passes with the table in LDS
    int force_lanes = 0;   // lanes per atom of the LDS-table force pass (0: as force_path)
    int cell_brick = 1;    // cells are numbered in bricks of cell_brick^3 (1: row-major)
    msm::LtShape lt{1024, 1, 1024, 1};  // block / workgroups per CU of the LDS-table passes
    int compact = 0;       // global-list path: force pass reads the density pass's in-cutoff list

    // adjoint scratch (allocated on first use)
    double *adj_dldf = nullptr, *adj_gbox = nullptr, *adj_g = nullptr, *adj_rho = nullptr,
           *adj_lam = nullptr, *adj_part = nullptr, *adj_out = nullptr;
    size_t adj_cap_n = 0, adj_cap_part = 0, adj_cap_out = 0, adj_cap_box = 0;

    // device-resident force matching (msm_set_spline_table, msm_fm_*): coefficients, targets
    // (caller order), scratch of the loss, the loss and the coefficient gradient
    double *fm_coef = nullptr, *fm_ft = nullptr, *fm_et = nullptr, *fm_part = nullptr,
           *fm_loss = nullptr, *fm_grad = nullptr;
    size_t fm_cap_coef = 0, fm_cap_ft = 0, fm_cap_et = 0, fm_cap_part = 0, fm_cap_loss = 0,
           fm_cap_grad = 0;
    double* fm_host = nullptr;  // pinned staging: coefficients in, then loss and gradient out
    hipEvent_t fm_ev = nullptr; // end of the last msm_set_spline_table on the stream
    // msm_fm_loss_grad's kernels and copies as graphs, by what they bake in (a few at most)
    std::map<std::vector<uint64_t>, hipGraphExec_t> fm_graphs;
    uint64_t graph_gen = 0;     // bumped by clear_graphs: every cached graph is stale
    bool fm_pending = false;    // host-synchronous table writers wait for fm_ev first
    size_t fm_cap_host = 0;
    int fm_n = -1;         // atom count the targets were given for (-1: none)
    double fm_w_e = 0.0;

    // LDS-staged list bricks
    int nbricks = 0;
    int nl_lanes = 1;  // lanes per home atom of the list build
    int nl_cap = 64;   // halo atoms staged per list-build workgroup
    int nl_wide = 0;   // bricks planned for the wide list kernel
    int list_sidx = 0; // the current list holds halo slots (written for the brick passes)
    int tuning = 0;    // msm_autotune is running: lists carry halo slots for every candidate
    std::vector<double> occ;  // mean atoms per non-empty cell, per box (at upload)
    int* d_brick_box = nullptr;

    // per-class kernel timing (MSM_OPT_KERNEL_TIMING)
    std::vector<hipEvent_t> evpool;
    std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> pending;
    double kms[MSM_KCLASS_COUNT] = {};
    long long kcnt[MSM_KCLASS_COUNT] = {};

    std::map<GraphKey, hipGraphExec_t> graphs;
};

namespace {

int set_err(msm_sim* s, int code, const std::string& msg) {
    if (s) s->err = msg;
    return code;
}

#define HIPCHK(sim, call)                                                                    \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_err(sim, MSM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(msm_sim* s, T** p, size_t count) {
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(s, MSM_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return MSM_OK;
}

template <typename T>
void dfree(T** p) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
}

// Device buffer of at least count elements: kept when it already holds that many (so refilling a
// table of the same size allocates nothing and keeps its address), else freed and reallocated.
template <typename T>
int dreserve(msm_sim* s, T** p, size_t count, size_t& cap) {
    if (*p && cap >= count) return MSM_OK;
    cap = 0;
    const int rc = dalloc(s, p, count);
    if (rc == MSM_OK) cap = count;
    return rc;
}

void clear_graphs(msm_sim* s) {
    for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
    s->graphs.clear();
    for (auto& kv : s->fm_graphs) (void)hipGraphExecDestroy(kv.second);
    s->fm_graphs.clear();
    ++s->graph_gen;
}

void free_atoms(msm_sim* s) {
    for (int k = 0; k < 2; ++k) {
        dfree(&s->pos[k]);
        dfree(&s->set[k].v);
        dfree(&s->set[k].f);
        dfree(&s->set[k].dis);
        dfree(&s->set[k].ityp);
        dfree(&s->set[k].statu);
        dfree(&s->set[k].boxid);
        dfree(&s->set[k].gid);
        dfree(&s->set[k].cell);
    }
    dfree(&s->epot);
    s->n = s->n_alloc = 0;
    s->has_atoms = false;
    s->has_list = false;
}

AtomsDev atoms_view(msm_sim* s, int set, int pcur) {
    AtomsDev a{};
    const size_t na = static_cast<size_t>(s->n_alloc);
    AtomSet& z = s->set[set];
    a.pos = s->pos[pcur];
    a.pos_out = s->pos[1 - pcur];
    a.vx = z.v; a.vy = z.v + na; a.vz = z.v + 2 * na;
    a.fx = z.f; a.fy = z.f + na; a.fz = z.f + 2 * na;
    a.dx = z.dis; a.dy = z.dis + na; a.dz = z.dis + 2 * na;
    a.ityp = z.ityp;
    a.statu = z.statu;
    a.boxid = z.boxid;
    a.epot = s->epot;
    a.n = s->n;
    return a;
}

bool use_lds_table(const msm_sim* s);

ListDev list_view(msm_sim* s) {
    ListDev l{};
    l.kvois = s->kvois;
    l.indi = s->indi;
    l.sidx = s->sidx;
    const bool lt = use_lds_table(s);
    const bool cmp = !lt && s->compact && s->force_path > 0 && s->force_path <= 4 &&
                     s->force_path != 3;
    l.cidx = cmp ? s->cidx : nullptr;
    l.ckv = cmp ? s->ckv : nullptr;
    l.cmask = (lt && s->compact) ? s->cmask : nullptr;
    l.stride = msm::kListRow;
    l.mx = s->mx;
    l.scap = (s->nlist_kind == 1 || !s->list_sidx) ? 0 : s->nl_wide ? s->nl_cap : msm::kNlCapPass;
    return l;
}

int check_ready(msm_sim* s, bool need_list) {
    if (!s) return MSM_ERR_ARG;
    if (!s->has_table) return set_err(s, MSM_ERR_STATE, "no force table (msm_set_force_table)");
    if (!s->has_ctrl) return set_err(s, MSM_ERR_STATE, "no control parameters (msm_set_ctrl)");
    if (!s->has_boxes) return set_err(s, MSM_ERR_STATE, "no boxes (msm_set_boxes)");
    if (!s->has_atoms) return set_err(s, MSM_ERR_STATE, "no atoms (msm_set_atoms)");
    if (s->tab_ngroup != s->ngroup)
        return set_err(s, MSM_ERR_ARG, "force table and control parameters disagree on NGROUP");
    if (need_list && !s->has_list)
        return set_err(s, MSM_ERR_STATE, "no neighbour list (msm_build_neighbors)");
    return MSM_OK;
}

// ---- optional per-launch timing with HIP events on the context's stream ----
hipEvent_t take_event(msm_sim* s) {
    if (!s->evpool.empty()) {
        hipEvent_t e = s->evpool.back();
        s->evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct TimedScope {
    msm_sim* s;
    int cls;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TimedScope(msm_sim* sim, int c) : s(sim), cls(c) {
        if (s->timing) {
            e0 = take_event(s);
            e1 = take_event(s);
            (void)hipEventRecord(e0, s->stream);
        }
    }
    ~TimedScope() {
        if (s->timing) {
            (void)hipEventRecord(e1, s->stream);
            s->pending.emplace_back(cls, e0, e1);
        }
    }
};

void resolve_timing(msm_sim* s) {
    for (auto& p : s->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, std::get<1>(p), std::get<2>(p)) == hipSuccess) {
            s->kms[std::get<0>(p)] += ms;
            s->kcnt[std::get<0>(p)] += 1;
        }
        s->evpool.push_back(std::get<1>(p));
        s->evpool.push_back(std::get<2>(p));
    }
    s->pending.clear();
}

// Device error flags: cleared with a fill kernel so the clear can sit inside captured graphs.
void clear_flags(msm_sim* s) { msm::launch_fill_int(s->d_err, 0, msm::kErrSlots, s->stream); }

int read_flags(msm_sim* s) {
    HIPCHK(s, hipMemcpyAsync(s->h_err, s->d_err, sizeof(int) * msm::kErrSlots,
                             hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    resolve_timing(s);
    if (s->h_err[msm::kErrOutOfBox])
        return set_err(s, MSM_ERR_OUTOFBOX, "an active atom lies outside its box's cell grid");
    if (s->h_err[msm::kErrOverflow])
        return set_err(s, MSM_ERR_NLIST_OVERFLOW,
                       "an atom has more neighbours than MXKVOIS (list truncated)");
    if (s->h_err[msm::kErrEmbd])
        return set_err(s, MSM_ERR_EMBD_RANGE, "a density fell beyond the embedding table");
    return MSM_OK;
}

// Neighbour-list rebuild on the stream (capturable: kernels only).  Reads the current set and
// position buffer, writes the other set and buffer, then flips both.
// keep_f: carry the forces into the new order (not needed inside msm_run, where a force pass
// follows every rebuild).
void enqueue_rebuild(msm_sim* s, bool keep_f = true) {
    TimedScope ts(s, MSM_KCLASS_REBUILD);
    const int src = s->scur, dst = 1 - s->scur;
    const int p0 = s->pcur, p1 = 1 - s->pcur;
    AtomsDev a_src = atoms_view(s, src, p0);
    AtomsDev a_dst = atoms_view(s, dst, p1);
    a_dst.pos = s->pos[p1];
    const int nc = s->ncell_total;
    msm::launch_fill_int(s->sb.count, 0, nc + 2, s->stream);
    msm::launch_cell_assign(a_src.pos, a_src.statu, a_src.boxid, s->d_boxes, s->n, nc, s->sb,
                            s->d_err, s->stream);
    msm::launch_scan_exclusive(s->sb.count, s->sb.start, nc + 2, s->sb.scan_tmp, s->stream);
    msm::launch_scatter_sort(s->n, nc, s->sb, s->set[src].gid, s->stream);
    msm::launch_gather(s->n, s->n_alloc, s->sb.perm, a_src, s->set[src].gid, a_dst,
                       s->set[dst].gid, s->sb.key, s->set[dst].cell, keep_f ? 1 : 0, s->stream);
    TimedScope tl(s, MSM_KCLASS_NLIST);
    if (s->nlist_kind == 1) {
        s->list_sidx = 0;
        msm::launch_nlist(s->n, s->pos[p1], s->set[dst].ityp, s->set[dst].statu,
                          s->set[dst].boxid, s->set[dst].cell, s->sb.start, s->d_boxes, s->ngroup,
                          s->d_rnl2, s->mx, s->kvois, s->indi, msm::kListRow, s->d_err, s->stream);
    } else {
        // halo slots only for the brick passes (a list without them sends those passes to
        // the global list until the next rebuild)
        s->list_sidx = s->force_path == 0 || s->force_path == 5 || s->tuning;
        msm::launch_nlist_bricks(s->nl_wide, s->nbricks, s->nl_lanes, s->nl_cap, s->list_sidx,
                                 s->d_brick_box, s->n, nc,
                                 s->pos[p1],
                                 s->set[dst].ityp, s->sb.start, s->d_boxes, s->ngroup, s->d_rnl2,
                                 s->mx, s->kvois, s->indi, s->sidx, msm::kListRow, s->d_err,
                                 s->stream);
    }
    s->scur = dst;
    s->pcur = p1;
}

bool multibox(const msm_sim* s) { return s->boxes.size() > 1; }

// Brick shape of the LDS-staged list build, per box, and lanes per home atom.  From the box's
// mean occupancy m of non-empty cells (binning the uploaded positions on the host), a brick of
// `home` cells with a `halo` of cells costs one workgroup about
//     ceil(home*m*T / kNlBlock) * 27*m/T   (candidate tests per lane, T lanes per atom)
//   + 3*halo*m / kNlBlock                  (staging)
// and serves home*m atoms; its LDS holds halo*(1.3 m + 1) atoms, which sets how many workgroups
// share a CU.  The largest box picks T in {1, 2, 4, 8}; every box then takes the brick (at most
// 4 cells a side) with the most atoms per cost per CU.  Bricks that overflow at run time fall
// back to the global-memory scan, so the choice affects speed only.
struct BrickChoice {
    int e[3] = {1, 1, 1};
    int lanes = 1;
    int cap = 0;
    double score = -1.0;
};

BrickChoice best_brick(const BoxDev& b, double m, int lanes_fixed, int ngroup) {
    BrickChoice best;
    const double est = 1.3 * m + 1.0;
    const double lds_cu = 163840.0, lds_static = 6144.0;
    for (int lanes : {1, 2, 4, 8}) {
        if (lanes_fixed > 0 && lanes != lanes_fixed) continue;
        for (int bz = 1; bz <= 4; ++bz)
            for (int by = 1; by <= 4; ++by)
                for (int bx = 1; bx <= 4; ++bx) {
                    const int e[3] = {std::min(bx, b.ncell[0]), std::min(by, b.ncell[1]),
                                      std::min(bz, b.ncell[2])};
                    const int halo = (e[0] + 2) * (e[1] + 2) * (e[2] + 2);
                    const int home = e[0] * e[1] * e[2];
                    const int cap = static_cast<int>(std::ceil(halo * est));
                    if (cap > msm::kNlCap) continue;
                    const double lds = static_cast<double>(msm::nlist_lds_bytes(cap, ngroup)) +
                                       lds_static;
                    const double wgs = std::min(2.0, std::floor(lds_cu / lds));
                    const double atoms = home * m;
                    const double rounds = std::ceil(atoms * lanes / msm::kNlBlock);
                    const double cost = std::max(rounds, 1.0) * 27.0 * m / lanes +
                                        3.0 * halo * m / msm::kNlBlock;
                    const double score = wgs * atoms / cost;
                    if (score > best.score * (1.0 + 1e-9)) {
                        best.score = score;
                        best.lanes = lanes;
                        best.cap = cap;
                        for (int k = 0; k < 3; ++k) best.e[k] = e[k];
                    }
                }
    }
    return best;
}

// Small bricks: the largest brick (at most 4 cells a side) whose halo and home atoms, at
// 1.3 m + 1 atoms per cell, fit the kBlock-thread kernel's kNlCapPass LDS slots and 4 atoms per
// thread.
void small_brick(const BoxDev& b, double m, int e_out[3]) {
    const double est = 1.3 * m + 1.0;
    int best_home = 0;
    e_out[0] = e_out[1] = e_out[2] = 1;
    for (int bz = 1; bz <= 4; ++bz)
        for (int by = 1; by <= 4; ++by)
            for (int bx = 1; bx <= 4; ++bx) {
                const int e[3] = {std::min(bx, b.ncell[0]), std::min(by, b.ncell[1]),
                                  std::min(bz, b.ncell[2])};
                const int halo = (e[0] + 2) * (e[1] + 2) * (e[2] + 2);
                const int home = e[0] * e[1] * e[2];
                if (halo * est > msm::kNlCapPass || home * est > 4 * msm::kBlock) continue;
                if (home > best_home || (home == best_home && e[0] > e_out[0])) {
                    best_home = home;
                    for (int k = 0; k < 3; ++k) e_out[k] = e[k];
                }
            }
}

// Bricks for the LDS list build of the current kind (s->nlist_kind; 0 picks wide bricks when
// the largest box holds 12 or more atoms per non-empty cell) from the occupancies of upload.
int replan_bricks(msm_sim* s) {
    if (s->occ.size() != s->boxes.size()) s->occ.assign(s->boxes.size(), 1.0);
    size_t largest = 0;
    for (size_t b = 0; b < s->boxes.size(); ++b)
        if (s->boxes[b].natoms > s->boxes[largest].natoms) largest = b;
    const int wide = s->nlist_kind == 3 || (s->nlist_kind == 0 && s->occ[largest] >= 12.0);
    std::vector<int> brick_box;
    int lanes = 1, cap = 64, off = 0;
    if (wide) lanes = best_brick(s->boxes[largest], s->occ[largest], 0, s->ngroup).lanes;
    for (size_t bidx = 0; bidx < s->boxes.size(); ++bidx) {
        BoxDev& b = s->boxes[bidx];
        int e[3];
        if (wide) {
            const BrickChoice c = best_brick(b, s->occ[bidx], lanes, s->ngroup);
            cap = std::max(cap, c.cap);
            for (int k = 0; k < 3; ++k) e[k] = c.e[k];
        } else {
            small_brick(b, s->occ[bidx], e);
        }
        int nb = 1;
        for (int k = 0; k < 3; ++k) {
            b.bdim[k] = e[k];
            b.nbrick[k] = (b.ncell[k] + e[k] - 1) / e[k];
            nb *= b.nbrick[k];
        }
        b.brick_off = off;
        for (int k = 0; k < nb; ++k) brick_box.push_back(static_cast<int>(bidx));
        off += nb;
    }
    s->nl_wide = wide;
    s->nl_lanes = lanes;
    s->nl_cap = std::min(cap, msm::kNlCap);
    s->nbricks = off;
    int rc = dalloc(s, &s->d_brick_box, brick_box.size());
    if (rc != MSM_OK) return rc;
    HIPCHK(s, hipMemcpy(s->d_brick_box, brick_box.data(), brick_box.size() * sizeof(int),
                        hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->d_boxes, s->boxes.data(), s->boxes.size() * sizeof(BoxDev),
                        hipMemcpyHostToDevice));
    return MSM_OK;
}

// Mean atoms per non-empty cell of every box, from the uploaded positions, then the brick plan.
int plan_bricks(msm_sim* s, const double* xp, int n) {
    s->occ.assign(s->boxes.size(), 1.0);
    int aoff = 0;
    for (size_t bidx = 0; bidx < s->boxes.size(); ++bidx) {
        BoxDev& b = s->boxes[bidx];
        const int nc = b.ncell[0] * b.ncell[1] * b.ncell[2];
        std::vector<int> cnt(nc, 0);
        for (int i = aoff; i < aoff + b.natoms; ++i) {
            int c3[3];
            for (int k = 0; k < 3; ++k) {
                const double x = xp[i + static_cast<size_t>(k) * n];
                int c = static_cast<int>((x - b.lo[k]) / b.len[k] * b.ncell[k]);
                c3[k] = std::min(std::max(c, 0), b.ncell[k] - 1);
            }
            ++cnt[c3[0] + b.ncell[0] * (c3[1] + b.ncell[1] * c3[2])];
        }
        aoff += b.natoms;
        long long occ = 0, nonempty = 0;
        for (int c : cnt)
            if (c > 0) {
                occ += c;
                ++nonempty;
            }
        s->occ[bidx] = nonempty ? static_cast<double>(occ) / nonempty : 1.0;
    }
    return replan_bricks(s);
}

// LDS-table passes: global list with 1, 2 or 4 lanes per atom or two lanes paired for 32-byte
// gathers (path 3), no compaction, and an LDS window that holds at least the upper half of every
// table kind in use.
// Lanes per atom of the LDS-table force pass: MSM_OPT_FORCE_LANES when set (not with lane pairs),
// else those of the density pass.
// Staged-brick passes (path 5): over a list from the brick list kernels that carries halo
// slots, with knot windows of at least a quarter of the table beside the staged halo.
bool use_staged(const msm_sim* s) {
    if (s->force_path != 5 || s->nlist_kind == 1 || !s->list_sidx) return false;
    const int scap = s->nl_wide ? s->nl_cap : msm::kNlCapPass;
    return msm::sb_plan(s->tab, s->lt, scap).ok;
}

// Path 5 without a usable staged plan runs the two-lane LDS-table passes.
int eff_path(const msm_sim* s) { return s->force_path == 5 ? 2 : s->force_path; }

int lt_force_lanes(const msm_sim* s) {
    return (s->force_lanes > 0 && eff_path(s) != 3) ? s->force_lanes : eff_path(s);
}

bool use_lds_table(const msm_sim* s) {
    if (!s->table_lds) return false;
    const int fp = eff_path(s);
    if (fp < 1 || fp > 4) return false;  // 3: lane-pair gathers
    // compaction: per-lane 64-bit masks of in-cutoff slots, written by the density pass for the
    // same lanes of the force pass (not with lane pairs, whose two lanes walk their slots together)
    if (s->compact && (fp == 3 || lt_force_lanes(s) != fp || (s->mx + fp - 1) / fp > 64))
        return false;
    const msm::LtWin w = msm::lt_window(s->tab, s->lt.wgs_d, s->lt.wgs_f);
    return 2 * w.wf >= s->tab.ntab2 && 2 * w.wd >= s->tab.ntab2;
}

// Brick passes with the table in LDS: path 0 over a list built by the brick kernels (its slots
// index the staged halos), with windows holding at least the top third of every kind in use
// beside a non-empty staged halo.
bool use_brick_table(const msm_sim* s) {
    if (!s->table_lds || s->force_path != 0 || s->nlist_kind == 1) return false;
    const msm::BtPlan p =
        msm::bt_plan(s->tab, s->lt, s->nl_wide ? s->nl_cap : msm::kNlCapPass);
    return 3 * p.w.wf >= s->tab.ntab2 && 3 * p.w.wd >= s->tab.ntab2 && p.cap_d > 0 &&
           p.cap_f > 0;
}

void enqueue_density(msm_sim* s) {
    TimedScope ts(s, MSM_KCLASS_DENSITY);
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    if (use_staged(s))
        msm::launch_density_sb(s->lt, s->nbricks, s->d_brick_box, a, list_view(s), s->tab,
                               s->d_boxes, s->sb.start, s->ncell_total, s->d_err, s->stream);
    else if (use_brick_table(s))
        msm::launch_density_bt(s->bt_lanes, s->lt, s->nbricks, s->d_brick_box, a, list_view(s),
                               s->tab, s->d_boxes, s->sb.start, s->ncell_total, s->d_err,
                               s->stream);
    else if (use_lds_table(s))
        msm::launch_density_lt(eff_path(s), s->lt, a, list_view(s), s->tab, s->boxes[0],
                               s->d_boxes, multibox(s), s->d_err, s->stream);
    else if (eff_path(s) == 3)
        msm::launch_density_pair(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s),
                                 s->d_err, s->stream);
    else if (eff_path(s) > 0)
        msm::launch_density_tpa(eff_path(s), a, list_view(s), s->tab, s->boxes[0], s->d_boxes,
                                multibox(s), s->d_err, s->stream);
    else if (s->nlist_kind == 1)
        msm::launch_density(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s),
                            s->d_err, s->stream);
    else
        msm::launch_density_bricks(s->nbricks, s->d_brick_box, a, list_view(s), s->tab,
                                   s->d_boxes, s->sb.start, s->ncell_total, s->d_err, s->stream);
}

// mode 0: force only; 1: force + corrector; 2: force + corrector + next predictor (flips pos).
void enqueue_force(msm_sim* s, int mode, int write_f, double h) {
    TimedScope ts(s, MSM_KCLASS_FORCE);
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    StepParams sp{};
    sp.h = h;
    sp.hs2 = h * 0.5;
    sp.h2s2 = h * sp.hs2;
    sp.mode = mode;
    sp.write_f = write_f;
    sp.track_dis = s->track_dis;
    if (use_staged(s))
        msm::launch_force_sb(s->lt, s->nbricks, s->d_brick_box, a, list_view(s), s->tab,
                             s->d_boxes, s->sb.start, s->ncell_total, sp, s->stream);
    else if (use_brick_table(s))
        msm::launch_force_bt(s->bt_lanes, s->lt, s->nbricks, s->d_brick_box, a, list_view(s),
                             s->tab, s->d_boxes, s->sb.start, s->ncell_total, sp, s->stream);
    else if (use_lds_table(s))
        msm::launch_force_lt(lt_force_lanes(s), s->lt, a, list_view(s), s->tab, s->boxes[0],
                             s->d_boxes, multibox(s), sp, s->stream);
    else if (eff_path(s) == 3)
        msm::launch_force_pair(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s), sp,
                               s->stream);
    else if (eff_path(s) > 0)
        msm::launch_force_tpa(eff_path(s), a, list_view(s), s->tab, s->boxes[0], s->d_boxes,
                              multibox(s), sp, s->stream);
    else if (s->nlist_kind == 1)
        msm::launch_force(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s), sp,
                          s->stream);
    else
        msm::launch_force_bricks(s->nbricks, s->d_brick_box, a, list_view(s), s->tab, s->d_boxes,
                                 s->sb.start, s->ncell_total, sp, s->stream);
    if (mode == 2) s->pcur = 1 - s->pcur;
}

// One segment of the step loop: [rebuild] then nsteps x (density, fused force); the final step
// of the run applies only the corrector and stores forces.
void enqueue_segment(msm_sim* s, int rebuild, int nsteps, int last, double h) {
    if (rebuild) enqueue_rebuild(s, false);
    for (int k = 0; k < nsteps; ++k) {
        enqueue_density(s);
        const bool final_step = last && (k == nsteps - 1);
        if (final_step)
            enqueue_force(s, 1, 1, h);
        else
            enqueue_force(s, 2, 0, h);
    }
}

// Host-side state change of a segment without enqueuing work (used when a cached graph replays
// the segment).
void advance_segment_state(msm_sim* s, int rebuild, int nsteps, int last) {
    if (rebuild) {
        s->scur = 1 - s->scur;
        s->pcur = 1 - s->pcur;
    }
    const int flips = last ? nsteps - 1 : nsteps;
    if (flips & 1) s->pcur = 1 - s->pcur;
}

int run_segment(msm_sim* s, int rebuild, int nsteps, int last, double h) {
    if (s->timing) {  // direct launches, each bracketed by events
        enqueue_segment(s, rebuild, nsteps, last, h);
        return MSM_OK;
    }
    uint64_t hbits = 0;
    std::memcpy(&hbits, &h, sizeof(hbits));
    GraphKey key{s->scur, s->pcur, rebuild, nsteps, last, s->track_dis, hbits};
    auto it = s->graphs.find(key);
    if (it == s->graphs.end()) {
        const int scur0 = s->scur, pcur0 = s->pcur;
        HIPCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        enqueue_segment(s, rebuild, nsteps, last, h);
        hipGraph_t g = nullptr;
        HIPCHK(s, hipStreamEndCapture(s->stream, &g));
        hipGraphExec_t ge = nullptr;
        hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess)
            return set_err(s, MSM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        it = s->graphs.emplace(key, ge).first;
        // the capture advanced the host state; rewind so the launch below replays from it
        s->scur = scur0;
        s->pcur = pcur0;
    }
    HIPCHK(s, hipGraphLaunch(it->second, s->stream));
    advance_segment_state(s, rebuild, nsteps, last);
    return MSM_OK;
}

// Workspace of msm_calc_ptensor_noalloc: per-box partial sums of the virial, in a fixed device
// array (each GPU that loads the library holds its own copy until the process ends) so that the
// call allocates nothing.  Host code must never pass `scratch` itself to a kernel (on the host it
// names a stand-in, not device memory): it uses the per-device address taken below.
constexpr size_t kScratchDoubles = size_t(1) << 19;  // 4 MiB
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchMaxDevices] = {};

// Once per device, with that device current: the device address of `scratch` on it.
int scratch_setup(int device) {
    if (device < 0 || device >= kScratchMaxDevices) return MSM_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return MSM_OK;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return MSM_ERR_HIP;
    g_scratch_ptr[device] = static_cast<double*>(ptr);
    return MSM_OK;
}

double* scratch_for(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kScratchMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

// ============================================================================================
// C-ABI
// ============================================================================================

extern "C" {

const char* msm_strerror(int code) {
    switch (code) {
        case MSM_OK: return "ok";
        case MSM_ERR_ARG: return "invalid argument";
        case MSM_ERR_HIP: return "HIP runtime error";
        case MSM_ERR_NOMEM: return "device allocation failed";
        case MSM_ERR_NLIST_OVERFLOW: return "neighbour list overflow";
        case MSM_ERR_OUTOFBOX: return "active atom out of box";
        case MSM_ERR_EMBD_RANGE: return "density beyond embedding table";
        case MSM_ERR_STATE: return "call out of order";
        default: return "unknown error";
    }
}

int msm_create(msm_sim** out, int device) {
    if (!out) return MSM_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MSM_ERR_HIP;
    if (device < 0 || device >= ndev) return MSM_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return MSM_ERR_HIP;
    const int src = scratch_setup(device);
    if (src != MSM_OK) return src;
    msm_sim* s = new msm_sim();
    s->device = device;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return MSM_ERR_HIP;
    }
    int rc = dalloc(s, &s->d_err, msm::kErrSlots);
    if (rc == MSM_OK) rc = dalloc(s, &s->d_stats, 2);
    if (rc == MSM_OK && hipHostMalloc(reinterpret_cast<void**>(&s->h_err),
                                      sizeof(int) * msm::kErrSlots) != hipSuccess)
        rc = MSM_ERR_NOMEM;
    if (rc != MSM_OK) {
        msm_destroy(s);
        return rc;
    }
    *out = s;
    return MSM_OK;
}

int msm_destroy(msm_sim* s) {
    if (!s) return MSM_OK;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    clear_graphs(s);
    free_atoms(s);
    for (double** p : {&s->adj_dldf, &s->adj_gbox, &s->adj_g, &s->adj_rho, &s->adj_lam,
                       &s->adj_part, &s->adj_out, &s->fm_coef, &s->fm_ft, &s->fm_et,
                       &s->fm_part, &s->fm_loss, &s->fm_grad})
        dfree(p);
    for (auto& p : s->pending) {
        (void)hipEventDestroy(std::get<1>(p));
        (void)hipEventDestroy(std::get<2>(p));
    }
    for (hipEvent_t e : s->evpool) (void)hipEventDestroy(e);
    if (s->fm_ev) (void)hipEventDestroy(s->fm_ev);
    dfree(&s->d_b2);
    dfree(&s->d_f4);
    dfree(&s->d_e4);
    dfree(&s->d_m4);
    dfree(&s->d_rnl2);
    dfree(&s->d_boxes);
    dfree(&s->d_brick_box);
    dfree(&s->sb.key);
    dfree(&s->sb.slot);
    dfree(&s->sb.count);
    dfree(&s->sb.start);
    dfree(&s->sb.perm);
    dfree(&s->sb.scan_tmp);
    dfree(&s->kvois);
    dfree(&s->indi);
    dfree(&s->sidx);
    dfree(&s->cidx);
    dfree(&s->ckv);
    dfree(&s->cmask);
    dfree(&s->d_err);
    dfree(&s->d_ebox);
    dfree(&s->d_stats);
    dfree(&s->d_boxwork);
    if (s->h_err) (void)hipHostFree(s->h_err);
    if (s->fm_host) (void)hipHostFree(s->fm_host);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return MSM_OK;
}

const char* msm_last_error(const msm_sim* s) { return s ? s->err.c_str() : "null context"; }

int msm_set_force_table(msm_sim* s, int pot_type, int ngroup, int nkind, int ntab, double csi,
                        const int* kpair, const double* potr, const double* fpotr,
                        const double* potb, const double* fpotb, int nkind1, int nembd,
                        double rhod, const int* kembd, const double* fembd,
                        const double* dfembd) {
    if (!s) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    // the copies below are not ordered after table kernels still queued on the stream
    if (s->fm_pending) {
        HIPCHK(s, hipEventSynchronize(s->fm_ev));
        s->fm_pending = false;
    }
    if (pot_type != MSM_FS_TYPE && pot_type != MSM_EAM_TYPE)
        return set_err(s, MSM_ERR_ARG, "pot_type must be MSM_FS_TYPE or MSM_EAM_TYPE");
    if (ngroup < 1 || ngroup > msm::kMaxGroup || nkind < 1 || ntab < 2 || !(csi > 0.0) ||
        !kpair || !potr || !fpotr || !potb || !fpotb)
        return set_err(s, MSM_ERR_ARG, "bad pair-table arguments");
    if (pot_type == MSM_EAM_TYPE &&
        (nkind1 < 1 || nembd < 2 || !(rhod > 0.0) || !kembd || !fembd || !dfembd))
        return set_err(s, MSM_ERR_ARG, "bad embedding-table arguments");
    const int ntab2 = ntab + 2;
    // kind-major padded knots with forward differences (v(i+1) - v(i) is the same IEEE
    // subtraction the reference evaluates per lookup, so precomputing it changes no bit)
    auto knot = [&](const double* T, int k, int i) -> double {
        if (i < 1 || i > ntab) return 0.0;
        return T[static_cast<size_t>(k) + static_cast<size_t>(i - 1) * nkind];
    };
    std::vector<double2> hb2(static_cast<size_t>(nkind) * ntab2);
    std::vector<double4> hf4(static_cast<size_t>(nkind) * ntab2), he4(hf4.size());
    for (int k = 0; k < nkind; ++k)
        for (int i = 0; i < ntab2; ++i) {
            const size_t a = static_cast<size_t>(k) * ntab2 + i;
            const double pb = knot(potb, k, i), pb1 = knot(potb, k, i + 1);
            hb2[a] = make_double2(pb, pb1 - pb);
            const double fr = knot(fpotr, k, i), fr1 = knot(fpotr, k, i + 1);
            const double fb = knot(fpotb, k, i), fb1 = knot(fpotb, k, i + 1);
            hf4[a] = make_double4(fr, fr1 - fr, fb, fb1 - fb);
            const double pr = knot(potr, k, i), pr1 = knot(potr, k, i + 1);
            he4[a] = make_double4(pr, pr1 - pr, pb, pb1 - pb);
        }
    TabDev t;
    std::memset(&t, 0, sizeof(t));  // byte-wise comparable with the current table
    t.pot_type = pot_type;
    t.ngroup = ngroup;
    t.ntab2 = ntab2;
    t.nkind = nkind;
    t.csi = csi;
    for (int i = 0; i < ngroup; ++i)
        for (int j = 0; j < ngroup; ++j) {
            const int kd = kpair[i + j * ngroup];
            if (kd < 1 || kd > nkind) return set_err(s, MSM_ERR_ARG, "KPAIR entry out of range");
            t.kpair[i * ngroup + j] = kd - 1;
        }
    if (pot_type == MSM_EAM_TYPE)
        for (int i = 0; i < ngroup; ++i)
            if (kembd[i] < 1 || kembd[i] > nkind1)
                return set_err(s, MSM_ERR_ARG, "KEMBD entry out of range");
    int rc = dreserve(s, &s->d_b2, hb2.size(), s->cap_b2);
    if (rc == MSM_OK) rc = dreserve(s, &s->d_f4, hf4.size(), s->cap_f4);
    if (rc == MSM_OK) rc = dreserve(s, &s->d_e4, he4.size(), s->cap_e4);
    if (rc != MSM_OK) return rc;
    HIPCHK(s, hipMemcpy(s->d_b2, hb2.data(), hb2.size() * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->d_f4, hf4.data(), hf4.size() * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->d_e4, he4.data(), he4.size() * sizeof(double4), hipMemcpyHostToDevice));
    if (pot_type == MSM_EAM_TYPE) {
        const int nembd2 = nembd + 2;
        auto eknot = [&](const double* T, int k, int i) -> double {
            if (i < 1 || i > nembd) return 0.0;
            return T[static_cast<size_t>(k) + static_cast<size_t>(i - 1) * nkind1];
        };
        std::vector<double4> hm4(static_cast<size_t>(nkind1) * nembd2);
        for (int k = 0; k < nkind1; ++k)
            for (int i = 0; i < nembd2; ++i) {
                const double f0 = eknot(fembd, k, i), f1 = eknot(fembd, k, i + 1);
                const double d0 = eknot(dfembd, k, i), d1 = eknot(dfembd, k, i + 1);
                hm4[static_cast<size_t>(k) * nembd2 + i] = make_double4(f0, f1 - f0, d0, d1 - d0);
            }
        for (int i = 0; i < ngroup; ++i) t.kembd[i] = kembd[i] - 1;
        rc = dreserve(s, &s->d_m4, hm4.size(), s->cap_m4);
        if (rc != MSM_OK) return rc;
        HIPCHK(s, hipMemcpy(s->d_m4, hm4.data(), hm4.size() * sizeof(double4), hipMemcpyHostToDevice));
        t.nembd = nembd;
        t.nembd2 = nembd2;
        t.nkind1 = nkind1;
        t.rhod = rhod;
    } else {
        dfree(&s->d_m4);
        t.nembd = 0;
        t.nembd2 = 0;
        t.nkind1 = 1;
        t.rhod = 1.0;
    }
    t.b2 = s->d_b2;
    t.f4 = s->d_f4;
    t.e4 = s->d_e4;
    t.m4 = s->d_m4;
    // keep control-dependent fields if the control block came first
    t.ru2max = s->tab.ru2max;
    for (int g = 0; g < msm::kMaxGroup; ++g) t.cm[g] = s->tab.cm[g];
    // captured segments hold the table by value: keep them only when nothing but knot values
    // changed (same buffers, sizes, kinds and grid)
    const bool same = s->has_table && s->tab_ngroup == ngroup &&
                      std::memcmp(&s->tab, &t, sizeof(t)) == 0;
    s->tab = t;
    s->tab_ngroup = ngroup;
    s->has_table = true;
    if (!same) clear_graphs(s);
    return MSM_OK;
}

int msm_set_ctrl(msm_sim* s, int ngroup, const double* ru, const double* nb_rm, int mxkvois,
                 const double* cm) {
    if (!s) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (ngroup < 1 || ngroup > msm::kMaxGroup || !ru || !nb_rm || !cm || mxkvois < 1)
        return set_err(s, MSM_ERR_ARG, "bad control arguments");
    s->ngroup = ngroup;
    s->mx = mxkvois;
    s->ru.assign(ru, ru + ngroup * ngroup);
    s->nb_rm.assign(nb_rm, nb_rm + ngroup * ngroup);
    s->cm.assign(cm, cm + ngroup);
    double ru2max = 0.0;
    std::vector<double> rnl2(static_cast<size_t>(ngroup) * ngroup);
    for (int i = 0; i < ngroup; ++i)
        for (int j = 0; j < ngroup; ++j) {
            const double r = ru[i + j * ngroup];
            ru2max = std::max(ru2max, r * r);
            const double q = nb_rm[i + j * ngroup];
            rnl2[static_cast<size_t>(i) * ngroup + j] = q * q;
        }
    for (int g = 0; g < ngroup; ++g) {
        if (!(cm[g] > 0.0)) return set_err(s, MSM_ERR_ARG, "masses must be positive");
        s->tab.cm[g] = cm[g];
    }
    s->tab.ru2max = ru2max;
    int rc = dalloc(s, &s->d_rnl2, rnl2.size());
    if (rc != MSM_OK) return rc;
    HIPCHK(s, hipMemcpy(s->d_rnl2, rnl2.data(), rnl2.size() * sizeof(double), hipMemcpyHostToDevice));
    s->has_ctrl = true;
    s->has_list = false;
    clear_graphs(s);
    return MSM_OK;
}

int msm_set_boxes(msm_sim* s, int nbox, const int* natoms, const double* lo, const double* len,
                  const int* ifpd) {
    if (!s) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (!s->has_ctrl) return set_err(s, MSM_ERR_STATE, "set control parameters before boxes");
    if (nbox < 1 || !natoms || !lo || !len || !ifpd)
        return set_err(s, MSM_ERR_ARG, "bad box arguments");
    double rmax = 0.0;
    for (double r : s->nb_rm) rmax = std::max(rmax, r);
    std::vector<BoxDev> bx(nbox);
    long long cells = 0;
    for (int b = 0; b < nbox; ++b) {
        BoxDev& d = bx[b];
        std::memset(&d, 0, sizeof(d));
        if (natoms[b] < 0) return set_err(s, MSM_ERR_ARG, "negative atom count");
        d.natoms = natoms[b];
        int nc = 1;
        for (int k = 0; k < 3; ++k) {
            d.lo[k] = lo[3 * b + k];
            d.len[k] = len[3 * b + k];
            if (!(d.len[k] > 0.0)) return set_err(s, MSM_ERR_ARG, "box length must be positive");
            d.hlen[k] = d.len[k] * 0.5;
            d.ifpd[k] = ifpd[3 * b + k] ? 1 : 0;
            d.ilen[k] = d.ifpd[k] ? 1.0 / d.len[k] : 0.0;
            // NCELL = max(3, int(L / max(NB_RM) - eps)), MD_NeighborsList.F90:474-483
            int c = static_cast<int>(d.len[k] / (1.0 * rmax) - 0.0001);
            if (c < 3) c = 3;
            d.ncell[k] = c;
            nc *= c;
        }
        d.cbs = s->cell_brick;
        d.cell_off = static_cast<int>(cells);
        cells += nc;
        if (cells > (1LL << 30)) return set_err(s, MSM_ERR_ARG, "too many cells");
    }
    int rc = dalloc(s, &s->d_boxes, bx.size());
    if (rc != MSM_OK) return rc;
    HIPCHK(s, hipMemcpy(s->d_boxes, bx.data(), bx.size() * sizeof(BoxDev), hipMemcpyHostToDevice));
    s->boxes = bx;
    s->ncell_total = static_cast<int>(cells);
    const int need = s->ncell_total + 2;
    if (need > s->sort_cap) {
        rc = dalloc(s, &s->sb.count, need);
        if (rc == MSM_OK) rc = dalloc(s, &s->sb.start, need);
        if (rc == MSM_OK) rc = dalloc(s, &s->sb.scan_tmp, need / 1024 + 2);
        if (rc != MSM_OK) return rc;
        s->sort_cap = need;
    }
    rc = dalloc(s, &s->d_ebox, nbox);
    if (rc != MSM_OK) return rc;
    s->has_boxes = true;
    s->has_atoms = false;
    s->has_list = false;
    clear_graphs(s);
    return MSM_OK;
}

int msm_set_atoms(msm_sim* s, int n, const double* xp, const double* xp1, const int* ityp,
                  const int* statu) {
    if (!s) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (!s->has_boxes) return set_err(s, MSM_ERR_STATE, "set boxes before atoms");
    if (n < 0 || (n > 0 && (!xp || !xp1 || !ityp || !statu)))
        return set_err(s, MSM_ERR_ARG, "bad atom arguments");
    long long tot = 0;
    for (const BoxDev& b : s->boxes) tot += b.natoms;
    if (tot != n) return set_err(s, MSM_ERR_ARG, "sum of box atom counts differs from n");
    for (int i = 0; i < n; ++i)
        if (ityp[i] < 1 || ityp[i] > s->ngroup) return set_err(s, MSM_ERR_ARG, "ITYP out of range");
    clear_graphs(s);
    free_atoms(s);
    const int na = std::max(64, (n + 63) / 64 * 64);
    int rc = MSM_OK;
    for (int k = 0; k < 2 && rc == MSM_OK; ++k) {
        rc = dalloc(s, &s->pos[k], na);
        if (rc == MSM_OK) rc = dalloc(s, &s->set[k].v, 3 * static_cast<size_t>(na));
        if (rc == MSM_OK) rc = dalloc(s, &s->set[k].f, 3 * static_cast<size_t>(na));
        if (rc == MSM_OK) rc = dalloc(s, &s->set[k].dis, 3 * static_cast<size_t>(na));
        if (rc == MSM_OK) rc = dalloc(s, &s->set[k].ityp, na);
        if (rc == MSM_OK) rc = dalloc(s, &s->set[k].statu, na);
        if (rc == MSM_OK) rc = dalloc(s, &s->set[k].boxid, na);
        if (rc == MSM_OK) rc = dalloc(s, &s->set[k].gid, na);
        if (rc == MSM_OK) rc = dalloc(s, &s->set[k].cell, na);
    }
    if (rc == MSM_OK) rc = dalloc(s, &s->epot, na);
    if (rc == MSM_OK) rc = dalloc(s, &s->sb.key, na);
    if (rc == MSM_OK) rc = dalloc(s, &s->sb.slot, na);
    if (rc == MSM_OK) rc = dalloc(s, &s->sb.perm, na);
    if (rc == MSM_OK) rc = dalloc(s, &s->kvois, na);
    const size_t icap = static_cast<size_t>(na) * s->mx;
    if (rc == MSM_OK) rc = dalloc(s, &s->indi, icap);
    // the staged passes issue a full chunk of slot loads past an atom's last slot
    if (rc == MSM_OK) rc = dalloc(s, &s->sidx, icap + msm::kSbSlack);
    if (rc == MSM_OK) rc = dalloc(s, &s->cidx, icap);
    if (rc == MSM_OK) rc = dalloc(s, &s->ckv, static_cast<size_t>(na) * 8);
    if (rc == MSM_OK) rc = dalloc(s, &s->cmask, static_cast<size_t>(na) * 4);
    if (rc != MSM_OK) {
        free_atoms(s);
        return rc;
    }
    s->indi_cap = icap;
    s->n = n;
    s->n_alloc = na;
    s->scur = 0;
    s->pcur = 0;
    std::vector<double4> hp(na, make_double4(0, 0, 0, 0));
    std::vector<double> hv(3 * static_cast<size_t>(na), 0.0);
    std::vector<int> ht(na, 0), hs(na, 0), hb(na, 0), hg(na, 0);
    for (int i = 0; i < n; ++i) {
        hp[i] = make_double4(xp[i], xp[i + n], xp[i + 2 * static_cast<size_t>(n)], 0.0);
        for (int k = 0; k < 3; ++k)
            hv[i + static_cast<size_t>(k) * na] = xp1[i + static_cast<size_t>(k) * n];
        ht[i] = ityp[i] - 1;
        hs[i] = statu[i];
        hg[i] = i;
    }
    int off = 0;
    for (size_t b = 0; b < s->boxes.size(); ++b) {
        for (int i = 0; i < s->boxes[b].natoms; ++i) hb[off + i] = static_cast<int>(b);
        off += s->boxes[b].natoms;
    }
    AtomSet& z = s->set[0];
    HIPCHK(s, hipMemcpy(s->pos[0], hp.data(), na * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(z.v, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice));
    std::fill(hv.begin(), hv.end(), 0.0);
    HIPCHK(s, hipMemcpy(z.f, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(z.dis, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(z.ityp, ht.data(), na * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(z.statu, hs.data(), na * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(z.boxid, hb.data(), na * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(z.gid, hg.data(), na * sizeof(int), hipMemcpyHostToDevice));
    std::vector<double> hz(na, 0.0);
    HIPCHK(s, hipMemcpy(s->epot, hz.data(), na * sizeof(double), hipMemcpyHostToDevice));
    rc = plan_bricks(s, xp, n);
    if (rc != MSM_OK) return rc;
    s->has_atoms = true;
    s->has_list = false;
    return MSM_OK;
}

int msm_get_atoms(msm_sim* s, double* xp, double* xp1, double* fp, int* statu, double* dis) {
    if (!s) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (!s->has_atoms) return set_err(s, MSM_ERR_STATE, "no atoms");
    const int n = s->n;
    const size_t na = s->n_alloc;
    AtomSet& z = s->set[s->scur];
    double* tmp = nullptr;
    int* itmp = nullptr;
    int rc = dalloc(s, &tmp, 3 * na);
    if (rc == MSM_OK) rc = dalloc(s, &itmp, na);
    if (rc != MSM_OK) {
        dfree(&tmp);
        return rc;
    }
    auto pull3 = [&](const double* src3, double* host) -> int {
        for (int k = 0; k < 3; ++k)
            msm::launch_unsort_d(src3 + k * na, tmp + k * static_cast<size_t>(n), z.gid, n, s->stream);
        HIPCHK(s, hipMemcpyAsync(host, tmp, 3 * static_cast<size_t>(n) * sizeof(double),
                                 hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        return MSM_OK;
    };
    if (xp && n > 0) {
        for (int k = 0; k < 3; ++k)
            msm::launch_unsort_pos(s->pos[s->pcur], tmp + k * static_cast<size_t>(n), z.gid, n, k, s->stream);
        HIPCHK(s, hipMemcpyAsync(xp, tmp, 3 * static_cast<size_t>(n) * sizeof(double),
                                 hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
    }
    if (rc == MSM_OK && xp1 && n > 0) rc = pull3(z.v, xp1);
    if (rc == MSM_OK && fp && n > 0) rc = pull3(z.f, fp);
    if (rc == MSM_OK && dis && n > 0) rc = pull3(z.dis, dis);
    if (rc == MSM_OK && statu && n > 0) {
        msm::launch_unsort_i(z.statu, itmp, z.gid, n, s->stream);
        hipError_t e = hipMemcpyAsync(statu, itmp, n * sizeof(int), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    }
    dfree(&tmp);
    dfree(&itmp);
    return rc;
}

static int pull_sorted_d(msm_sim* s, const double* src, int comp_pos, double* host) {
    const int n = s->n;
    if (n == 0) return MSM_OK;
    double* tmp = nullptr;
    int rc = dalloc(s, &tmp, n);
    if (rc != MSM_OK) return rc;
    AtomSet& z = s->set[s->scur];
    if (comp_pos >= 0)
        msm::launch_unsort_pos(s->pos[s->pcur], tmp, z.gid, n, comp_pos, s->stream);
    else
        msm::launch_unsort_d(src, tmp, z.gid, n, s->stream);
    hipError_t e = hipMemcpyAsync(host, tmp, n * sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    dfree(&tmp);
    if (e != hipSuccess) return set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    return MSM_OK;
}

int msm_put_atoms(msm_sim* s, const double* xp, const double* xp1) {
    if (!s) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (!s->has_atoms) return set_err(s, MSM_ERR_STATE, "no atoms");
    const int n = s->n;
    if (n == 0 || (!xp && !xp1)) return MSM_OK;
    const size_t na = s->n_alloc;
    AtomSet& z = s->set[s->scur];
    double* tmp = nullptr;
    int rc = dalloc(s, &tmp, 3 * static_cast<size_t>(n));
    if (rc != MSM_OK) return rc;
    hipError_t e = hipSuccess;
    if (xp) {
        e = hipMemcpyAsync(tmp, xp, 3 * static_cast<size_t>(n) * sizeof(double),
                           hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) msm::launch_sort_pos(tmp, s->pos[s->pcur], z.gid, n, s->stream);
    }
    if (e == hipSuccess && xp1) {
        if (xp) e = hipStreamSynchronize(s->stream);  // tmp is reused
        if (e == hipSuccess)
            e = hipMemcpyAsync(tmp, xp1, 3 * static_cast<size_t>(n) * sizeof(double),
                               hipMemcpyHostToDevice, s->stream);
        for (int k = 0; e == hipSuccess && k < 3; ++k)
            msm::launch_sort_d(tmp + k * static_cast<size_t>(n), z.v + k * na, z.gid, n,
                               s->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    dfree(&tmp);
    if (e != hipSuccess) return set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    return MSM_OK;
}

// Caller-order arrays of ncomp components (stride n) added into sorted per-atom device arrays
// dst[k] (k < ncomp); src is a host or device pointer.
static int add_caller_order(msm_sim* s, const double* src, int on_device, double* const* dst,
                            int ncomp, const char* what) {
    if (!s) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (!s->has_atoms) return set_err(s, MSM_ERR_STATE, "no atoms");
    if (!src) return set_err(s, MSM_ERR_ARG, std::string(what) + ": null array");
    const int n = s->n;
    if (n == 0) return MSM_OK;
    const size_t cnt = static_cast<size_t>(ncomp) * n;
    double* tmp = nullptr;
    const double* d = src;
    if (!on_device) {
        int rc = dalloc(s, &tmp, cnt);
        if (rc != MSM_OK) return rc;
        hipError_t e = hipMemcpyAsync(tmp, src, cnt * sizeof(double), hipMemcpyHostToDevice,
                                      s->stream);
        if (e != hipSuccess) {
            dfree(&tmp);
            return set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
        }
        d = tmp;
    }
    const int* gid = s->set[s->scur].gid;
    for (int k = 0; k < ncomp; ++k)
        msm::launch_add_sorted(d + static_cast<size_t>(k) * n, dst[k], gid, n, s->stream);
    hipError_t e = hipStreamSynchronize(s->stream);
    dfree(&tmp);
    if (e != hipSuccess) return set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    return MSM_OK;
}

int msm_add_forces(msm_sim* s, const double* fext, int on_device) {
    if (!s) return MSM_ERR_ARG;
    const size_t na = s->n_alloc;
    AtomSet& z = s->set[s->scur];
    double* const dst[3] = {z.f, z.f + na, z.f + 2 * na};
    return add_caller_order(s, fext, on_device, dst, 3, "msm_add_forces");
}

int msm_add_epot(msm_sim* s, const double* eext, int on_device) {
    if (!s) return MSM_ERR_ARG;
    double* const dst[1] = {s->epot};
    int rc = add_caller_order(s, eext, on_device, dst, 1, "msm_add_epot");
    if (rc != MSM_OK || s->n == 0 || !s->has_list) return rc;
    // keep the box totals of msm_get_box_epot in step with the per-atom energies
    msm::launch_box_sum(s->epot, s->sb.start, s->d_boxes, static_cast<int>(s->boxes.size()),
                        s->d_ebox, s->stream);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MSM_OK;
}

int msm_put_dis(msm_sim* s, const double* dis) {
    if (!s) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (!s->has_atoms) return set_err(s, MSM_ERR_STATE, "no atoms");
    if (!dis) return set_err(s, MSM_ERR_ARG, "msm_put_dis: null array");
    const int n = s->n;
    if (n == 0) return MSM_OK;
    const size_t na = s->n_alloc;
    AtomSet& z = s->set[s->scur];
    double* tmp = nullptr;
    int rc = dalloc(s, &tmp, 3 * static_cast<size_t>(n));
    if (rc != MSM_OK) return rc;
    hipError_t e = hipMemcpyAsync(tmp, dis, 3 * static_cast<size_t>(n) * sizeof(double),
                                  hipMemcpyHostToDevice, s->stream);
    for (int k = 0; e == hipSuccess && k < 3; ++k)
        msm::launch_sort_d(tmp + k * static_cast<size_t>(n), z.dis + k * na, z.gid, n, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    dfree(&tmp);
    if (e != hipSuccess) return set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    return MSM_OK;
}

int msm_get_den(msm_sim* s, double* den) {
    if (!s || !den) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (!s->has_atoms) return set_err(s, MSM_ERR_STATE, "no atoms");
    return pull_sorted_d(s, nullptr, 3, den);
}

int msm_get_epot(msm_sim* s, double* epot) {
    if (!s || !epot) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (!s->has_atoms) return set_err(s, MSM_ERR_STATE, "no atoms");
    return pull_sorted_d(s, s->epot, -1, epot);
}

int msm_get_box_epot(msm_sim* s, double* ebox) {
    if (!s || !ebox) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (!s->has_list) return set_err(s, MSM_ERR_STATE, "no neighbour list");
    msm::launch_box_sum(s->epot, s->sb.start, s->d_boxes, static_cast<int>(s->boxes.size()),
                        s->d_ebox, s->stream);
    HIPCHK(s, hipMemcpyAsync(ebox, s->d_ebox, s->boxes.size() * sizeof(double),
                             hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MSM_OK;
}

int msm_build_neighbors(msm_sim* s) {
    int rc = check_ready(s, false);
    if (rc != MSM_OK) return rc;
    (void)hipSetDevice(s->device);
    clear_flags(s);
    enqueue_rebuild(s);
    s->has_list = true;
    return read_flags(s);
}

int msm_get_neighbors(msm_sim* s, int* kvois, int* indi, int ld) {
    if (!s || !kvois || !indi) return MSM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (!s->has_list) return set_err(s, MSM_ERR_STATE, "no neighbour list");
    const int n = s->n;
    if (ld < n) return set_err(s, MSM_ERR_ARG, "ld smaller than the atom count");
    if (n == 0) return MSM_OK;
    int *kv = nullptr, *ix = nullptr;
    int rc = dalloc(s, &kv, n);
    if (rc == MSM_OK) rc = dalloc(s, &ix, static_cast<size_t>(ld) * s->mx);
    if (rc != MSM_OK) {
        dfree(&kv);
        return rc;
    }
    msm::launch_unsort_list(s->kvois, s->indi, msm::kListRow, s->set[s->scur].gid, n, s->mx, kv, ix,
                            ld, s->stream);
    hipError_t e = hipMemcpyAsync(kvois, kv, n * sizeof(int), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(indi, ix, static_cast<size_t>(ld) * s->mx * sizeof(int),
                           hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    dfree(&kv);
    dfree(&ix);
    if (e != hipSuccess) return set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    return MSM_OK;
}

int msm_calc_force(msm_sim* s) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    (void)hipSetDevice(s->device);
    clear_flags(s);
    enqueue_density(s);
    enqueue_force(s, 0, 1, 0.0);
    return read_flags(s);
}

int msm_calc_density(msm_sim* s) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    (void)hipSetDevice(s->device);
    clear_flags(s);
    enqueue_density(s);
    return read_flags(s);
}

int msm_calc_epot(msm_sim* s) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    (void)hipSetDevice(s->device);
    clear_flags(s);
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    msm::launch_epot(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s), s->d_err,
                     s->stream);
    return read_flags(s);
}

int msm_calc_ptensor(msm_sim* s, double* vtensor) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (!vtensor) return set_err(s, MSM_ERR_ARG, "null argument");
    (void)hipSetDevice(s->device);
    const int nbox = static_cast<int>(s->boxes.size());
    double* vir = nullptr;
    double* sums = nullptr;
    rc = dalloc(s, &vir, 6 * static_cast<size_t>(std::max(s->n, 1)));
    if (rc == MSM_OK) rc = dalloc(s, &sums, 6 * static_cast<size_t>(nbox));
    if (rc != MSM_OK) {
        dfree(&vir);
        return rc;
    }
    clear_flags(s);
    enqueue_density(s);
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    msm::launch_virial(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s), vir,
                       s->stream);
    for (int c = 0; c < 6; ++c)
        msm::launch_box_sum(vir + static_cast<size_t>(c) * s->n, s->sb.start, s->d_boxes, nbox,
                            sums + static_cast<size_t>(c) * nbox, s->stream);
    std::vector<double> h(6 * static_cast<size_t>(nbox));
    hipError_t e = hipMemcpyAsync(h.data(), sums, h.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    dfree(&vir);
    dfree(&sums);
    if (e != hipSuccess) return set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    static const int row[6] = {0, 1, 2, 0, 0, 1}, col[6] = {0, 1, 2, 1, 2, 2};
    for (int b = 0; b < nbox; ++b)
        for (int c = 0; c < 6; ++c) {
            const double v = h[static_cast<size_t>(c) * nbox + b];
            vtensor[9 * b + row[c] + 3 * col[c]] = v;
            vtensor[9 * b + col[c] + 3 * row[c]] = v;
        }
    return read_flags(s);
}

int msm_calc_ptensor_noalloc(msm_sim* s, double* vtensor) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (!vtensor) return set_err(s, MSM_ERR_ARG, "null argument");
    (void)hipSetDevice(s->device);
    double* ws = scratch_for(s->device);
    if (!ws) return set_err(s, MSM_ERR_STATE, "no workspace address for this device");
    const int nbox = static_cast<int>(s->boxes.size());
    int most = 1;
    for (const BoxDev& b : s->boxes) most = std::max(most, b.natoms);
    const size_t nchunk = (static_cast<size_t>(most) + msm::kBlock - 1) / msm::kBlock;
    const size_t need = static_cast<size_t>(nbox) * nchunk * 6 + 9 * static_cast<size_t>(nbox);
    if (need > kScratchDoubles)
        return set_err(s, MSM_ERR_ARG,
                       "msm_calc_ptensor_noalloc: nbox * ceil(largest box's atoms / 256) * 6 + "
                       "9 * nbox = " + std::to_string(need) + " doubles exceeds its fixed " +
                       std::to_string(kScratchDoubles) + "-double workspace");
    double* part = ws;
    double* out = ws + static_cast<size_t>(nbox) * nchunk * 6;
    clear_flags(s);
    enqueue_density(s);
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    msm::launch_virial_boxes(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s),
                             s->sb.start, nbox, s->ncell_total, static_cast<int>(nchunk), part,
                             out, s->stream);
    HIPCHK(s, hipMemcpyAsync(vtensor, out, 9 * static_cast<size_t>(nbox) * sizeof(double),
                             hipMemcpyDeviceToHost, s->stream));
    return read_flags(s);
}

int msm_calc_atomic_stress(msm_sim* s, double* avp) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (!avp) return set_err(s, MSM_ERR_ARG, "null argument");
    (void)hipSetDevice(s->device);
    const int n = s->n;
    if (n == 0) return MSM_OK;
    double* vir = nullptr;
    double* out = nullptr;
    rc = dalloc(s, &vir, 6 * static_cast<size_t>(n));
    if (rc == MSM_OK) rc = dalloc(s, &out, 9 * static_cast<size_t>(n));
    if (rc != MSM_OK) {
        dfree(&vir);
        return rc;
    }
    clear_flags(s);
    enqueue_density(s);
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    msm::launch_virial(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s), vir,
                       s->stream);
    msm::launch_unsort_stress(vir, out, s->set[s->scur].gid, n, s->stream);
    hipError_t e = hipMemcpyAsync(avp, out, 9 * static_cast<size_t>(n) * sizeof(double),
                                  hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    dfree(&vir);
    dfree(&out);
    if (e != hipSuccess) return set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    return read_flags(s);
}

// Per-box sum and count of the kinetic energies that count (CALEKIN_KERNEL); ek (device, n) keeps
// the per-atom values.
static int box_kinetic(msm_sim* s, double* ek, std::vector<double>& sum, std::vector<double>& cnt) {
    const int nbox = static_cast<int>(s->boxes.size());
    double* d = nullptr;
    int rc = dalloc(s, &d, 2 * static_cast<size_t>(nbox));
    if (rc != MSM_OK) return rc;
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    msm::launch_ekin(a, s->tab, ek, s->stream);
    msm::launch_box_ekin(ek, s->sb.start, s->d_boxes, nbox, d, d + nbox, s->stream);
    std::vector<double> h(2 * static_cast<size_t>(nbox));
    hipError_t e = hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    dfree(&d);
    if (e != hipSuccess) return set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    sum.assign(h.begin(), h.begin() + nbox);
    cnt.assign(h.begin() + nbox, h.end());
    return MSM_OK;
}

int msm_damping(msm_sim* s) {
    int rc = check_ready(s, false);
    if (rc != MSM_OK) return rc;
    (void)hipSetDevice(s->device);
    msm::launch_damping(atoms_view(s, s->scur, s->pcur), s->stream);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MSM_OK;
}

int msm_check_timestep(msm_sim* s, double th, double mxd2, int* exceeded) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (!exceeded || !(th > 0.0) || !(mxd2 >= 0.0))
        return set_err(s, MSM_ERR_ARG, "th > 0, mxd2 >= 0 and an output flag are required");
    (void)hipSetDevice(s->device);
    // the context's flag slot (no allocation per call: the step search calls this repeatedly)
    int* flag = s->d_err + msm::kErrDt;
    msm::launch_fill_int(flag, 0, 1, s->stream);
    msm::launch_check_dt(atoms_view(s, s->scur, s->pcur), s->tab, th, mxd2, flag, s->stream);
    HIPCHK(s, hipMemcpyAsync(s->h_err + msm::kErrDt, flag, sizeof(int), hipMemcpyDeviceToHost,
                             s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    *exceeded = s->h_err[msm::kErrDt] ? 1 : 0;
    return MSM_OK;
}

int msm_thermalize(msm_sim* s, double temp, double kb, unsigned long long seed) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (!(temp >= 0.0) || !(kb > 0.0))
        return set_err(s, MSM_ERR_ARG, "temperature >= 0 and kb > 0 are required");
    (void)hipSetDevice(s->device);
    const size_t need = 4 * s->boxes.size();
    if (s->boxwork_cap < need) {
        dfree(&s->d_boxwork);
        s->boxwork_cap = 0;
        if ((rc = dalloc(s, &s->d_boxwork, need)) != MSM_OK) return rc;
        s->boxwork_cap = need;
    }
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    msm::launch_thermalize(a, s->tab, s->set[s->scur].gid, temp * kb, seed, s->draws++,
                           s->sb.start, s->d_boxes, static_cast<int>(s->boxes.size()),
                           s->ncell_total, s->d_boxwork, s->stream);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MSM_OK;
}

int msm_calc_ekin(msm_sim* s, double kb, double* ekin, double* tbox) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (!(kb > 0.0)) return set_err(s, MSM_ERR_ARG, "kb must be positive");
    (void)hipSetDevice(s->device);
    double* ek = nullptr;
    if ((rc = dalloc(s, &ek, std::max(s->n, 1))) != MSM_OK) return rc;
    std::vector<double> sum, cnt;
    rc = box_kinetic(s, ek, sum, cnt);
    if (rc == MSM_OK && ekin) rc = pull_sorted_d(s, ek, -1, ekin);
    dfree(&ek);
    if (rc != MSM_OK) return rc;
    if (tbox)
        for (size_t b = 0; b < sum.size(); ++b)
            tbox[b] = cnt[b] > 0.0 ? 2.0 * (sum[b] / cnt[b]) / (3.0 * kb) : 0.0;
    return MSM_OK;
}

int msm_scale_velocities(msm_sim* s, double temp, double kb) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (!(kb > 0.0) || !(temp >= 0.0)) return set_err(s, MSM_ERR_ARG, "kb > 0 and temp >= 0");
    (void)hipSetDevice(s->device);
    const int nbox = static_cast<int>(s->boxes.size());
    double* ek = nullptr;
    double* scal = nullptr;
    rc = dalloc(s, &ek, std::max(s->n, 1));
    if (rc == MSM_OK) rc = dalloc(s, &scal, nbox);
    std::vector<double> sum, cnt;
    if (rc == MSM_OK) rc = box_kinetic(s, ek, sum, cnt);
    std::vector<double> f(nbox);
    for (int b = 0; rc == MSM_OK && b < nbox; ++b) {
        const double mean = cnt[b] > 0.0 ? sum[b] / cnt[b] : 0.0;
        if (!(mean > 0.0))
            rc = set_err(s, MSM_ERR_ARG, "current temperature of a box is zero in scaling velocity");
        else
            f[b] = std::sqrt(temp * 3.0 * kb * 0.5 / mean);  // VelScaling_DEV (:1390-1448)
    }
    if (rc == MSM_OK) {
        hipError_t e = hipMemcpyAsync(scal, f.data(), nbox * sizeof(double),
                                      hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) {
            msm::launch_velscale(atoms_view(s, s->scur, s->pcur), scal, s->stream);
            e = hipStreamSynchronize(s->stream);
        }
        if (e != hipSuccess) rc = set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    }
    dfree(&ek);
    dfree(&scal);
    return rc;
}

int msm_dyn_damp(msm_sim* s, int maxsteps, double mindele, double h, int* steps,
                 double* delepot) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (maxsteps < 0) return set_err(s, MSM_ERR_ARG, "maxsteps must be >= 0");
    (void)hipSetDevice(s->device);
    const int n = s->n;
    double* epot0 = nullptr;
    unsigned long long* dmax = nullptr;
    rc = dalloc(s, &epot0, std::max(n, 1));
    if (rc == MSM_OK) rc = dalloc(s, &dmax, 1);
    if (rc != MSM_OK) {
        dfree(&epot0);
        return rc;
    }
    clear_flags(s);
    // UpdateEPOT_ForceClass, then copy (Do_DynDamp_Forsteps_DEV, MD_DiffScheme_GPU.F90:1809-1883)
    msm::launch_epot(atoms_view(s, s->scur, s->pcur), list_view(s), s->tab, s->boxes[0],
                     s->d_boxes, multibox(s), s->d_err, s->stream);
    hipError_t e = hipMemcpyAsync(epot0, s->epot, std::max(n, 1) * sizeof(double),
                                  hipMemcpyDeviceToDevice, s->stream);
    int it = 1;
    double del = 0.0;
    for (; e == hipSuccess && it <= maxsteps; ++it) {
        AtomsDev a = atoms_view(s, s->scur, s->pcur);
        msm::launch_damping(a, s->stream);
        msm::launch_predictor(a, s->tab, s->boxes[0], s->d_boxes, multibox(s), h, s->stream);
        enqueue_density(s);
        enqueue_force(s, 0, 1, 0.0);
        a = atoms_view(s, s->scur, s->pcur);
        msm::launch_epot(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s), s->d_err,
                         s->stream);
        unsigned long long bits = 0;
        e = hipMemsetAsync(dmax, 0, sizeof(unsigned long long), s->stream);
        if (e != hipSuccess) break;
        msm::launch_epot_maxdelta(s->epot, epot0, n, dmax, s->stream);
        e = hipMemcpyAsync(&bits, dmax, sizeof(bits), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) break;
        std::memcpy(&del, &bits, sizeof(del));
        if (del <= mindele) break;  // converged: the reference leaves before the corrector
        msm::launch_correction(a, s->tab, h, s->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    dfree(&epot0);
    dfree(&dmax);
    if (e != hipSuccess) return set_err(s, MSM_ERR_HIP, hipGetErrorString(e));
    if (steps) *steps = it;  // maxsteps + 1 when not converged, as the reference's ITER
    if (delepot) *delepot = del;
    return read_flags(s);
}

// Secant step to the zero of the force projection on the search direction, from projections pf0
// (before the last step) and pf1 (after it), capped at maxdis; equal projections (a division by
// zero in the reference) take the largest step in the downhill direction.
static double secant_step(double step, double pf0, double pf1, double maxdis) {
    double st = pf1 != pf0 ? -step * pf1 / (pf1 - pf0) : (-step * pf1 >= 0.0 ? maxdis : -maxdis);
    if (std::fabs(st) > maxdis) st = st > 0.0 ? maxdis : -maxdis;
    return st;
}

namespace {

// Device scratch of one quench call (freed on every exit) and the small steps it is
// built from; every value the host branches on is read back with a stream synchronisation.
struct Quench {
    msm_sim* s;
    int n = 0;
    double* pf = nullptr;     // previous forces (dPreFP) or the unit direction (dDirect)
    double* dx = nullptr;     // displacement of the next move (dDXP)
    double* w = nullptr;      // forces of the last direction update (dF0, conjugate gradients)
    double* epot0 = nullptr;  // per-atom energy of the reference configuration (dEPOT0)
    double* part = nullptr;   // partial sums of the reductions
    double* out = nullptr;    // reduced values
    unsigned long long* bits = nullptr;
    hipError_t e = hipSuccess;

    explicit Quench(msm_sim* sim) : s(sim), n(sim->n) {}
    ~Quench() {
        dfree(&pf);
        dfree(&dx);
        dfree(&w);
        dfree(&epot0);
        dfree(&part);
        dfree(&out);
        dfree(&bits);
    }
    int alloc() {
        const size_t m = static_cast<size_t>(std::max(n, 1));
        int rc = dalloc(s, &pf, 3 * m);
        if (rc == MSM_OK) rc = dalloc(s, &dx, 3 * m);
        if (rc == MSM_OK) rc = dalloc(s, &w, 3 * m);
        if (rc == MSM_OK) rc = dalloc(s, &epot0, m);
        if (rc == MSM_OK) rc = dalloc(s, &part, 2 * static_cast<size_t>(msm::kQuenchBlocks));
        if (rc == MSM_OK) rc = dalloc(s, &out, 2);
        if (rc == MSM_OK) rc = dalloc(s, &bits, 1);
        return rc;
    }
    msm::P3 vec(double* base) const { return {base, base + n, base + 2 * static_cast<size_t>(n)}; }
    msm::P3 forces() const {
        AtomsDev a = atoms_view(s, s->scur, s->pcur);
        return {a.fx, a.fy, a.fz};
    }
    bool ok() const { return e == hipSuccess; }
    // CalForce_ForceClass: density and force passes on the current positions, forces stored
    void force() {
        enqueue_density(s);
        enqueue_force(s, 0, 1, 0.0);
    }
    // UpdateEPOT_ForceClass
    void epot() {
        msm::launch_epot(atoms_view(s, s->scur, s->pcur), list_view(s), s->tab, s->boxes[0],
                         s->d_boxes, multibox(s), s->d_err, s->stream);
    }
    void copy_epot0() {
        if (ok())
            e = hipMemcpyAsync(epot0, s->epot, std::max(n, 1) * sizeof(double),
                               hipMemcpyDeviceToDevice, s->stream);
    }
    double read_bits() {
        unsigned long long b = 0;
        if (ok()) e = hipMemcpyAsync(&b, bits, sizeof(b), hipMemcpyDeviceToHost, s->stream);
        if (ok()) e = hipStreamSynchronize(s->stream);
        double v = 0.0;
        std::memcpy(&v, &b, sizeof(v));
        return ok() ? v : 0.0;
    }
    // DevMaxAbsval of a 3-component vector
    double maxabs(msm::P3 v) {
        if (ok()) e = hipMemsetAsync(bits, 0, sizeof(unsigned long long), s->stream);
        if (ok()) msm::launch_maxabs3(v, n, bits, s->stream);
        return read_bits();
    }
    // max |EPOT - EPOT0| over atoms, then EPOT0 = EPOT
    double maxdelta() {
        if (ok()) e = hipMemsetAsync(bits, 0, sizeof(unsigned long long), s->stream);
        if (ok()) msm::launch_epot_maxdelta(s->epot, epot0, n, bits, s->stream);
        return read_bits();
    }
    void read_out(double* h, int k) {
        if (ok()) e = hipMemcpyAsync(h, out, k * sizeof(double), hipMemcpyDeviceToHost, s->stream);
        if (ok()) e = hipStreamSynchronize(s->stream);
    }
    double dot(msm::P3 u, msm::P3 v) {
        double r = 0.0;
        if (ok()) msm::launch_dot3(u, v, n, part, out, s->stream);
        read_out(&r, 1);
        return r;
    }
    // sum of u . (v - w)
    double dot_diff(msm::P3 u, msm::P3 v, msm::P3 wv) {
        double r[2] = {0.0, 0.0};
        if (ok()) msm::launch_dxdf(u, v, wv, n, part, out, s->stream);
        read_out(r, 2);
        return r[0];
    }
    void scale(msm::P3 src, double c, msm::P3 dst) {
        if (ok()) msm::launch_scale3(src, c, n, dst, s->stream);
    }
    void move(msm::P3 d) {
        if (ok())
            msm::launch_move3(atoms_view(s, s->scur, s->pcur), d, s->boxes[0], s->d_boxes,
                              multibox(s), s->stream);
    }
};

}  // namespace

int msm_steepest(msm_sim* s, int maxsteps, int linesearch, double alpha, double maxdis,
                 double mindis, double mindele, int* steps, double* delepot, double* maxmove) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (maxsteps < 0 || !(maxdis > 0.0) || !(mindis >= 0.0) || !(alpha > 0.0))
        return set_err(s, MSM_ERR_ARG, "maxsteps >= 0, maxdis > 0, mindis >= 0 and alpha > 0");
    (void)hipSetDevice(s->device);
    Quench q(s);
    if ((rc = q.alloc()) != MSM_OK) return rc;
    clear_flags(s);
    const msm::P3 DX = q.vec(q.dx), PF = q.vec(q.pf);
    int it = 0;
    double del = 0.0, mm = 0.0;
    if (!linesearch) {
        // Do_Steepest0_Forsteps_DEV (MD_SteepestScheme_GPU.F90:20-152)
        q.force();
        q.scale(q.forces(), 1.0, PF);
        q.scale(q.forces(), alpha, DX);
        mm = q.maxabs(DX);
        bool done = false;
        if (mm > maxdis)
            q.scale(DX, maxdis / mm, DX);
        else if (mm <= mindis)
            done = true;  // converged at the first step: nothing moves
        if (!done) {
            q.epot();
            q.copy_epot0();
            q.move(DX);
            int iflag = 0;
            for (int i = 1; q.ok() && i <= maxsteps; ++i) {
                q.force();
                double d2[2] = {0.0, 0.0};
                if (q.ok()) msm::launch_dxdf(DX, PF, q.forces(), q.n, q.part, q.out, s->stream);
                q.read_out(d2, 2);
                // ALPHA = DOTDXDF / DOTDF, replaced by STEEPEST_Alpha when negative (also when the
                // forces did not change, where the reference divides by zero)
                double al = d2[1] > 0.0 ? d2[0] / d2[1] : -1.0;
                if (al < 0.0) al = alpha;
                q.scale(q.forces(), al, DX);
                mm = q.maxabs(DX);
                if (mm > maxdis) q.scale(DX, maxdis / mm, DX);
                q.move(DX);
                if (mm <= mindis) {
                    iflag = i;
                    break;
                }
                q.epot();
                del = q.maxdelta();
                if (del <= mindele) {
                    iflag = i;
                    break;
                }
                q.scale(q.forces(), 1.0, PF);
            }
            it = iflag ? iflag : maxsteps + 1;
        }
    } else {
        // Do_Steepest1_Forsteps_DEV (MD_SteepestScheme_GPU.F90:156-260): line search along the
        // normalised force
        const msm::P3 D = PF;
        double step = maxdis;
        while (q.ok() && it <= maxsteps) {
            q.force();
            q.epot();
            q.copy_epot0();
            const double nrm = q.dot(q.forces(), q.forces());
            if (!(nrm > 0.0)) break;  // no force: already at a stationary point
            q.scale(q.forces(), 1.0 / std::sqrt(nrm), D);
            step = maxdis;
            q.scale(D, step, DX);
            double pf0 = q.dot(q.forces(), D);
            while (q.ok() && it <= maxsteps) {
                q.move(DX);
                q.force();
                const double pf1 = q.dot(q.forces(), D);
                step = secant_step(step, pf0, pf1, maxdis);
                q.scale(D, step, DX);
                ++it;
                if (std::fabs(step) <= mindis) break;
                pf0 = pf1;
            }
            q.epot();
            del = q.maxdelta();
            mm = std::fabs(step);
            if (del <= mindele) break;
        }
    }
    if (q.ok()) q.e = hipStreamSynchronize(s->stream);
    if (!q.ok()) return set_err(s, MSM_ERR_HIP, hipGetErrorString(q.e));
    if (steps) *steps = it;
    if (delepot) *delepot = del;
    if (maxmove) *maxmove = mm;
    return read_flags(s);
}

int msm_cg(msm_sim* s, int maxsteps, int linesearch, double maxdis, double mindis, double mindele,
           int* steps, double* delepot) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (maxsteps < 0 || !(maxdis > 0.0) || !(mindis >= 0.0))
        return set_err(s, MSM_ERR_ARG, "maxsteps >= 0, maxdis > 0 and mindis >= 0");
    (void)hipSetDevice(s->device);
    Quench q(s);
    if ((rc = q.alloc()) != MSM_OK) return rc;
    clear_flags(s);
    const double eps = 1.0e-64;  // EPS of Do_CG0/1_Forsteps_DEV
    const msm::P3 DIR = q.vec(q.pf), F0 = q.vec(q.w), DXP = q.vec(q.dx);
    int it = 0;
    double del = 0.0;
    // set-up of Do_CG0/1_Forsteps_DEV (MD_CGScheme_GPU.F90:16-291): direction = force
    q.force();
    q.epot();
    q.copy_epot0();
    q.scale(q.forces(), 1.0, DIR);
    q.scale(q.forces(), 1.0, F0);
    double f0n = q.dot(F0, F0);
    double pf0 = q.dot(q.forces(), DIR);
    // Polak-Ribiere update after a line minimisation: d = F + (F . (F - F0) / |F0|^2) d, F0 = F
    auto new_direction = [&]() {
        q.force();
        const double mf1 = q.dot_diff(q.forces(), q.forces(), F0);
        if (q.ok()) msm::launch_dir_update(q.forces(), mf1 / f0n, q.n, DIR, s->stream);
        q.scale(q.forces(), 1.0, F0);
        f0n = q.dot(F0, F0);
        pf0 = q.dot(q.forces(), DIR);
    };
    if (f0n > eps) {
        if (!linesearch) {
            // Do_CG0_Forsteps_DEV: one trial step and one secant step per direction
            int iflag = 0;
            for (int i = 1; q.ok() && i <= maxsteps; ++i) {
                double step = maxdis / std::sqrt(f0n);
                q.scale(DIR, step, DXP);
                q.move(DXP);
                q.force();
                const double pf1 = q.dot(q.forces(), DIR);
                step = secant_step(step, pf0, pf1, maxdis);
                q.scale(DIR, step / std::sqrt(f0n), DXP);
                q.move(DXP);
                q.epot();
                del = q.maxdelta();
                if (del <= mindele) {
                    iflag = i;
                    break;
                }
                new_direction();
                if (f0n <= eps) {
                    iflag = i;
                    break;
                }
            }
            it = iflag ? iflag : maxsteps + 1;
        } else {
            // Do_CG1_Forsteps_DEV: secant line search to |step| <= mindis along each direction
            while (q.ok() && it <= maxsteps) {
                double step = maxdis / std::sqrt(f0n);
                q.scale(DIR, step, DXP);
                q.move(DXP);
                q.force();
                double pf1 = q.dot(q.forces(), DIR);
                ++it;
                while (q.ok() && it <= maxsteps) {
                    step = secant_step(step, pf0, pf1, maxdis);
                    q.scale(DIR, step / std::sqrt(f0n), DXP);
                    q.move(DXP);
                    ++it;
                    if (std::fabs(step) <= mindis) break;
                    pf0 = pf1;
                    q.force();
                    pf1 = q.dot(q.forces(), DIR);
                }
                q.epot();
                del = q.maxdelta();
                if (del <= mindele) break;
                new_direction();
                if (f0n <= eps) break;
            }
        }
    }
    if (q.ok()) q.e = hipStreamSynchronize(s->stream);
    if (!q.ok()) return set_err(s, MSM_ERR_HIP, hipGetErrorString(q.e));
    if (steps) *steps = it;
    if (delepot) *delepot = del;
    return read_flags(s);
}

int msm_predictor(msm_sim* s, double h) {
    int rc = check_ready(s, false);
    if (rc != MSM_OK) return rc;
    (void)hipSetDevice(s->device);
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    msm::launch_predictor(a, s->tab, s->boxes[0], s->d_boxes, multibox(s), h, s->stream);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MSM_OK;
}

int msm_correction(msm_sim* s, double h) {
    int rc = check_ready(s, false);
    if (rc != MSM_OK) return rc;
    (void)hipSetDevice(s->device);
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    msm::launch_correction(a, s->tab, h, s->stream);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MSM_OK;
}

int msm_run(msm_sim* s, int nsteps, int nb_uptab, int step0, double h) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (nsteps < 0) return set_err(s, MSM_ERR_ARG, "nsteps must be >= 0");
    if (nsteps == 0) return MSM_OK;
    (void)hipSetDevice(s->device);
    clear_flags(s);
    // predictor of the first step (stand-alone kernel; later predictors run fused)
    {
        TimedScope ts(s, MSM_KCLASS_PREDICTOR);
        AtomsDev a = atoms_view(s, s->scur, s->pcur);
        msm::launch_predictor(a, s->tab, s->boxes[0], s->d_boxes, multibox(s), h, s->stream);
    }
    const int first = step0 + 1, last = step0 + nsteps;
    int sstep = first;
    while (sstep <= last) {
        const int rebuild = (nb_uptab > 0 && sstep % nb_uptab == 0) ? 1 : 0;
        // the segment runs until the step before the next rebuild or the end of the run
        int end = last;
        if (nb_uptab > 0) {
            int next = (sstep / nb_uptab + 1) * nb_uptab;
            end = std::min(last, next - 1);
        }
        const int cnt = end - sstep + 1;
        rc = run_segment(s, rebuild, cnt, end == last ? 1 : 0, h);
        if (rc != MSM_OK) return rc;
        sstep = end + 1;
    }
    return read_flags(s);
}

int msm_sync(msm_sim* s) {
    if (!s) return MSM_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MSM_OK;
}

int msm_set_option(msm_sim* s, int option, int value) {
    if (!s) return MSM_ERR_ARG;
    switch (option) {
        case MSM_OPT_TRACK_DIS:
            s->track_dis = value ? 1 : 0;
            return MSM_OK;
        case MSM_OPT_KERNEL_TIMING:
            s->timing = value ? 1 : 0;
            return MSM_OK;
        case MSM_OPT_NLIST_KERNEL: {
            if (value < 0 || value > 3)
                return set_err(s, MSM_ERR_ARG, "list kernel must be 0, 1, 2 or 3");
            s->nlist_kind = value;
            clear_graphs(s);
            if (!s->boxes.empty() && s->has_atoms) {
                const int rc = replan_bricks(s);
                if (rc != MSM_OK) return rc;
                s->has_list = false;  // brick slots of the list follow the plan: rebuild first
            }
            return MSM_OK;
        }
        case MSM_OPT_FORCE_PATH:
            if (value < 0 || value > 8 || (value > 5 && value != 8))
                return set_err(s, MSM_ERR_ARG, "force path must be 0, 1, 2, 3, 4, 5 or 8");
            s->force_path = value;
            clear_graphs(s);
            return MSM_OK;
        case MSM_OPT_COMPACT_LIST:
            s->compact = value ? 1 : 0;
            clear_graphs(s);
            return MSM_OK;
        case MSM_OPT_CHUNK_SWEEP:
            s->lt.sweep = value ? 1 : 0;
            clear_graphs(s);
            return MSM_OK;
        case MSM_OPT_TABLE_LDS:
            s->table_lds = value ? 1 : 0;
            clear_graphs(s);
            return MSM_OK;
        case MSM_OPT_CELL_BRICK:
            if (value < 1 || value > 16)
                return set_err(s, MSM_ERR_ARG, "cell brick must be 1..16");
            s->cell_brick = value;
            for (auto& b : s->boxes) b.cbs = value;
            if (!s->boxes.empty())
                HIPCHK(s, hipMemcpy(s->d_boxes, s->boxes.data(), s->boxes.size() * sizeof(BoxDev),
                                    hipMemcpyHostToDevice));
            s->has_list = false;  // the sorted order follows the numbering: rebuild first
            clear_graphs(s);
            return MSM_OK;
        case MSM_OPT_BRICK_LANES:
            if (value != 1 && value != 2 && value != 4)
                return set_err(s, MSM_ERR_ARG, "brick lanes must be 1, 2 or 4");
            s->bt_lanes = value;
            clear_graphs(s);
            return MSM_OK;
        case MSM_OPT_FORCE_LANES:
            if (value != 0 && value != 1 && value != 2 && value != 4)
                return set_err(s, MSM_ERR_ARG, "force lanes must be 0, 1, 2 or 4");
            s->force_lanes = value;
            clear_graphs(s);
            return MSM_OK;
        case MSM_OPT_LT_DENSITY_BLOCK:
        case MSM_OPT_LT_FORCE_BLOCK:
            if (value != 256 && value != 512 && value != 1024)
                return set_err(s, MSM_ERR_ARG, "LDS-table block must be 256, 512 or 1024");
            (option == MSM_OPT_LT_DENSITY_BLOCK ? s->lt.block_d : s->lt.block_f) = value;
            clear_graphs(s);
            return MSM_OK;
        case MSM_OPT_LT_DENSITY_WGS:
        case MSM_OPT_LT_FORCE_WGS:
            if (value < 1 || value > 16)
                return set_err(s, MSM_ERR_ARG, "LDS-table workgroups per CU must be 1..16");
            (option == MSM_OPT_LT_DENSITY_WGS ? s->lt.wgs_d : s->lt.wgs_f) = value;
            clear_graphs(s);
            return MSM_OK;
        default:
            return set_err(s, MSM_ERR_ARG, "unknown option");
    }
}

// The density/force kernels enqueue_density / enqueue_force launch.
static int pass_kernel(const msm_sim* s) {
    if (use_staged(s)) return MSM_PASS_STAGED;
    if (use_brick_table(s)) return MSM_PASS_BRICK_TABLE;
    if (use_lds_table(s)) return MSM_PASS_LDS_TABLE;
    if (eff_path(s) == 3) return MSM_PASS_LANE_PAIR;
    if (eff_path(s) > 0) return MSM_PASS_LANES;
    if (s->nlist_kind == 1) return MSM_PASS_GLOBAL;
    return MSM_PASS_BRICKS;
}

int msm_get_option(msm_sim* s, int option, int* value) {
    if (!s || !value) return MSM_ERR_ARG;
    switch (option) {
        case MSM_OPT_TRACK_DIS: *value = s->track_dis; return MSM_OK;
        case MSM_OPT_KERNEL_TIMING: *value = s->timing; return MSM_OK;
        case MSM_OPT_NLIST_KERNEL: *value = s->nlist_kind; return MSM_OK;
        case MSM_OPT_FORCE_PATH: *value = s->force_path; return MSM_OK;
        case MSM_OPT_COMPACT_LIST: *value = s->compact; return MSM_OK;
        case MSM_OPT_TABLE_LDS: *value = s->table_lds; return MSM_OK;
        case MSM_OPT_LT_DENSITY_BLOCK: *value = s->lt.block_d; return MSM_OK;
        case MSM_OPT_LT_DENSITY_WGS: *value = s->lt.wgs_d; return MSM_OK;
        case MSM_OPT_LT_FORCE_BLOCK: *value = s->lt.block_f; return MSM_OK;
        case MSM_OPT_LT_FORCE_WGS: *value = s->lt.wgs_f; return MSM_OK;
        case MSM_OPT_CELL_BRICK: *value = s->cell_brick; return MSM_OK;
        case MSM_OPT_BRICK_LANES: *value = s->bt_lanes; return MSM_OK;
        case MSM_OPT_FORCE_LANES: *value = s->force_lanes; return MSM_OK;
        case MSM_OPT_CHUNK_SWEEP: *value = s->lt.sweep; return MSM_OK;
        case MSM_OPT_PASS_KERNEL:
            if (!s->has_table) return set_err(s, MSM_ERR_STATE, "no force table");
            *value = pass_kernel(s);
            return MSM_OK;
        default:
            return set_err(s, MSM_ERR_ARG, "unknown option");
    }
}

int msm_kernel_times(msm_sim* s, double* ms, long long* count, int nclass, int reset) {
    if (!s || nclass < 0) return MSM_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    resolve_timing(s);
    for (int c = 0; c < nclass && c < MSM_KCLASS_COUNT; ++c) {
        if (ms) ms[c] = s->kms[c];
        if (count) count[c] = s->kcnt[c];
    }
    if (reset)
        for (int c = 0; c < MSM_KCLASS_COUNT; ++c) {
            s->kms[c] = 0.0;
            s->kcnt[c] = 0;
        }
    return MSM_OK;
}

int msm_autotune(msm_sim* s, int reps, int* choice, double* best_ms) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (reps < 1 || reps > 1000) return set_err(s, MSM_ERR_ARG, "reps must be 1..1000");
    (void)hipSetDevice(s->device);
    struct Cand {
        int fp, lds, bd, wd, bf, wf, lanes, flanes = 0;
    };
    std::vector<Cand> cands;
    if (msm::lt_window(s->tab, 1, 1).wf * 2 >= s->tab.ntab2) {
        for (int fp : {2, 3, 1, 4}) {
            cands.push_back({fp, 1, 1024, 1, 1024, 1, 1});
            cands.push_back({fp, 1, 256, 4, 256, 4, 1});
        }
        cands.push_back({2, 1, 512, 2, 512, 2, 1});
        cands.push_back({2, 1, 1024, 1, 1024, 1, 1, 1});  // force pass with one lane per atom
    }
    for (int fp : {2, 1, 3, 4, 8, 0}) cands.push_back({fp, 0, s->lt.block_d, s->lt.wgs_d,
                                                       s->lt.block_f, s->lt.wgs_f, 1});
    const bool brick_list = s->nlist_kind != 1;  // the list-kernel choice below keeps this
    const int saved_timing = s->timing;
    s->timing = 0;
    s->compact = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(s, hipEventCreate(&e0));
    HIPCHK(s, hipEventCreate(&e1));
    s->tuning = 1;  // cleared below on every path after this point
    auto time_it = [&](auto&& body, float& ms) -> int {
        body();  // warm-up (code object load, occupancy query)
        HIPCHK(s, hipEventRecord(e0, s->stream));
        for (int r = 0; r < reps; ++r) body();
        HIPCHK(s, hipEventRecord(e1, s->stream));
        HIPCHK(s, hipEventSynchronize(e1));
        HIPCHK(s, hipEventElapsedTime(&ms, e0, e1));
        ms /= static_cast<float>(reps);
        return MSM_OK;
    };
    clear_flags(s);
    // list-build kernel (a rebuild of unchanged positions reproduces the same order and list)
    float best_nl = 0.f;
    int nl_choice = s->nlist_kind;
    for (int g : {2, 3, 1}) {
        s->nlist_kind = g;
        if ((rc = replan_bricks(s)) != MSM_OK) break;
        float ms = 0.f;
        rc = time_it([&] { enqueue_rebuild(s); }, ms);
        if (rc != MSM_OK) break;
        if (g == 2 || ms < best_nl) {
            best_nl = ms;
            nl_choice = g;
        }
    }
    s->nlist_kind = nl_choice;
    if (rc == MSM_OK) rc = replan_bricks(s);
    if (rc == MSM_OK) enqueue_rebuild(s);
    // staged-brick passes (slot list of the brick list kernels; their workgroup shape is fixed,
    // the LDS-table shape fields are not used)
    if (brick_list && s->nlist_kind != 1)
        for (const Cand& c : {Cand{5, 1, 512, 2, 512, 2, 1}}) {
            s->lt = msm::LtShape{c.bd, c.wd, c.bf, c.wf};
            s->force_path = 5;
            s->table_lds = 1;
            if (use_staged(s)) cands.push_back(c);
        }
    // brick passes with the table in LDS (they read the slot list of the brick list kernels)
    if (brick_list && s->nlist_kind != 1)
        for (const Cand& c : {Cand{0, 1, 512, 2, 512, 2, 1}, Cand{0, 1, 1024, 1, 1024, 1, 2},
                              Cand{0, 1, 512, 2, 512, 2, 2}, Cand{0, 1, 1024, 1, 1024, 1, 1},
                              Cand{0, 1, 256, 4, 256, 4, 1}}) {
            s->lt = msm::LtShape{c.bd, c.wd, c.bf, c.wf};
            s->force_path = 0;
            s->table_lds = 1;
            if (use_brick_table(s)) cands.push_back(c);
        }
    // density + force passes
    float best = 0.f;
    size_t bi = 0;
    for (size_t c = 0; c < cands.size() && rc == MSM_OK; ++c) {
        s->force_path = cands[c].fp;
        s->table_lds = cands[c].lds;
        s->lt = msm::LtShape{cands[c].bd, cands[c].wd, cands[c].bf, cands[c].wf};
        s->bt_lanes = cands[c].lanes;
        s->force_lanes = cands[c].flanes;
        float ms = 0.f;
        rc = time_it([&] {
            enqueue_density(s);
            enqueue_force(s, 0, 1, 0.0);
        }, ms);
        if (rc == MSM_OK && (c == 0 || ms < best)) {
            best = ms;
            bi = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s->timing = saved_timing;
    s->tuning = 0;
    if (rc != MSM_OK) return rc;
    const Cand& w = cands[bi];
    s->force_path = w.fp;
    s->table_lds = w.lds;
    s->lt = msm::LtShape{w.bd, w.wd, w.bf, w.wf};
    s->bt_lanes = w.lanes;
    s->force_lanes = w.flanes;
    clear_graphs(s);
    // leave F' and F as the chosen kernels compute them
    enqueue_density(s);
    enqueue_force(s, 0, 1, 0.0);
    if (choice) {
        const int c[8] = {w.fp, w.lds, w.bd, w.wd, w.bf, w.wf, s->nlist_kind, w.lanes};
        for (int k = 0; k < 8; ++k) choice[k] = c[k];
    }
    if (best_ms) {
        best_ms[0] = best;
        best_ms[1] = best_nl;
    }
    return read_flags(s);
}

// Scratch of the adjoint passes for the current atoms and table (kept between calls); nb1, nb2:
// the bin counts of the pair and density passes, stride: the per-workgroup bin stride.
static int adj_reserve(msm_sim* s, size_t& nb1, size_t& nb2, size_t& stride) {
    int rc = MSM_OK;
    const int n = s->n, nbox = static_cast<int>(s->boxes.size());
    int nbin1 = 0, nbin2 = 0;
    msm::adjoint_bins(s->tab, nbin1, nbin2);
    nb1 = static_cast<size_t>(nbin1);
    nb2 = static_cast<size_t>(nbin2);
    stride = std::max(nb1, nb2);
    const size_t npart = stride * msm::adjoint_workgroups(n);
    if (s->adj_cap_n < static_cast<size_t>(s->n_alloc)) {
        dfree(&s->adj_dldf);
        dfree(&s->adj_g);
        dfree(&s->adj_rho);
        dfree(&s->adj_lam);
        s->adj_cap_n = 0;
        rc = dalloc(s, &s->adj_dldf, 3 * static_cast<size_t>(s->n_alloc));
        if (rc == MSM_OK) rc = dalloc(s, &s->adj_g, 4 * static_cast<size_t>(s->n_alloc));
        if (rc == MSM_OK) rc = dalloc(s, &s->adj_rho, s->n_alloc);
        if (rc == MSM_OK) rc = dalloc(s, &s->adj_lam, s->n_alloc);
        if (rc != MSM_OK) return rc;
        s->adj_cap_n = s->n_alloc;
    }
    if (s->adj_cap_part < npart) {
        dfree(&s->adj_part);
        s->adj_cap_part = 0;
        if ((rc = dalloc(s, &s->adj_part, npart)) != MSM_OK) return rc;
        s->adj_cap_part = npart;
    }
    if (s->adj_cap_out < nb1 + nb2) {
        dfree(&s->adj_out);
        s->adj_cap_out = 0;
        if ((rc = dalloc(s, &s->adj_out, nb1 + nb2)) != MSM_OK) return rc;
        s->adj_cap_out = nb1 + nb2;
    }
    if (s->adj_cap_box < static_cast<size_t>(nbox)) {
        dfree(&s->adj_gbox);
        s->adj_cap_box = 0;
        if ((rc = dalloc(s, &s->adj_gbox, nbox)) != MSM_OK) return rc;
        s->adj_cap_box = nbox;
    }
    return MSM_OK;
}

// The adjoint passes on dL/dF (s->adj_dldf, caller order) and dL/dE (s->adj_gbox) into the bins
// s->adj_out (pairs) and s->adj_out + nb1 (density).
static void adj_launch(msm_sim* s, size_t nb1, size_t stride, bool rho_ready = false) {
    msm::AdjDev d{};
    d.g = s->adj_g;
    d.gbox = s->adj_gbox;
    d.rho = s->adj_rho;
    d.lam = s->adj_lam;
    d.part = s->adj_part;
    d.part_stride = stride;
    AtomsDev a = atoms_view(s, s->scur, s->pcur);
    TimedScope ts(s, MSM_KCLASS_ADJOINT);
    msm::launch_adjoint(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s),
                        s->adj_dldf, s->set[s->scur].gid, d, s->adj_out, s->adj_out + nb1,
                        s->stream, rho_ready);
}

int msm_adjoint(msm_sim* s, const double* dldf, const double* dlde, double* g_potr,
                double* g_fpotr, double* g_potb, double* g_fpotb, double* g_fembd,
                double* g_dfembd) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (!dldf || !dlde || !g_potr || !g_fpotr || !g_potb || !g_fpotb)
        return set_err(s, MSM_ERR_ARG, "null argument");
    const bool eam = s->tab.pot_type != 0;
    if (eam && (!g_fembd || !g_dfembd))
        return set_err(s, MSM_ERR_ARG, "EAM tables need the embedding gradient arrays");
    (void)hipSetDevice(s->device);
    const int n = s->n, nbox = static_cast<int>(s->boxes.size());
    const int nt = s->tab.ntab2, ne = eam ? s->tab.nembd2 : 0;
    size_t nb1 = 0, nb2 = 0, stride = 0;
    if ((rc = adj_reserve(s, nb1, nb2, stride)) != MSM_OK) return rc;
    HIPCHK(s, hipMemcpyAsync(s->adj_dldf, dldf, 3 * static_cast<size_t>(n) * sizeof(double),
                             hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipMemcpyAsync(s->adj_gbox, dlde, nbox * sizeof(double), hipMemcpyHostToDevice,
                             s->stream));
    adj_launch(s, nb1, stride);
    std::vector<double> h(nb1 + nb2);
    HIPCHK(s, hipMemcpyAsync(h.data(), s->adj_out, h.size() * sizeof(double),
                             hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    resolve_timing(s);
    // bin (row r, knot i) -> the caller's [kind + (i-1)*NKIND]; with one atom type the kernels
    // keep one row, that of KPAIR(1,1) (KEMBD(1) for the embedding tables)
    const bool mt = s->tab.ngroup > 1;
    const int nk = s->tab.nkind, nk1 = eam ? s->tab.nkind1 : 1;
    const int rows = mt ? nk : 1, erows = mt ? nk1 : 1;
    const int ntab = nt - 2;
    for (int k = 0; k < nk; ++k)
        for (int i = 1; i <= ntab; ++i) {
            const size_t o = static_cast<size_t>(k) + static_cast<size_t>(i - 1) * nk;
            g_fpotr[o] = g_fpotb[o] = g_potr[o] = g_potb[o] = 0.0;
        }
    const size_t nkt = static_cast<size_t>(rows) * nt;
    for (int r = 0; r < rows; ++r) {
        const int k = mt ? r : s->tab.kpair[0];
        for (int i = 1; i <= ntab; ++i) {
            const size_t o = static_cast<size_t>(k) + static_cast<size_t>(i - 1) * nk;
            const size_t b = static_cast<size_t>(r) * nt + i;
            g_fpotr[o] = h[b];
            g_fpotb[o] = h[nkt + b];
            g_potr[o] = h[2 * nkt + b];
            g_potb[o] = h[nb1 + b];
        }
    }
    if (eam) {
        const int nembd = ne - 2;
        const size_t nke = static_cast<size_t>(erows) * ne;
        for (int k = 0; k < nk1; ++k)
            for (int i = 1; i <= nembd; ++i) {
                const size_t o = static_cast<size_t>(k) + static_cast<size_t>(i - 1) * nk1;
                g_fembd[o] = g_dfembd[o] = 0.0;
            }
        for (int r = 0; r < erows; ++r) {
            const int k = mt ? r : s->tab.kembd[0];
            for (int i = 1; i <= nembd; ++i) {
                const size_t o = static_cast<size_t>(k) + static_cast<size_t>(i - 1) * nk1;
                const size_t b = static_cast<size_t>(r) * ne + i;
                g_fembd[o] = h[nb1 + nkt + b];
                g_dfembd[o] = h[nb1 + nkt + nke + b];
            }
        }
    }
    return MSM_OK;
}

// Reserves *p for at least m doubles (capacity in cap).
static int fm_reserve(msm_sim* s, double** p, size_t m, size_t& cap) {
    if (cap >= m && *p) return MSM_OK;
    dfree(p);
    cap = 0;
    int rc = dalloc(s, p, m);
    if (rc == MSM_OK) cap = m;
    return rc;
}

// Pinned host staging of at least m doubles (PCIe copies without the runtime's pageable bounce).
static int fm_reserve_host(msm_sim* s, size_t m) {
    if (s->fm_cap_host >= m && s->fm_host) return MSM_OK;
    // a queued table build may still be copying from the old buffer
    if (s->fm_pending) {
        HIPCHK(s, hipEventSynchronize(s->fm_ev));
        s->fm_pending = false;
    }
    if (s->fm_host) (void)hipHostFree(s->fm_host);
    s->fm_host = nullptr;
    s->fm_cap_host = 0;
    if (hipHostMalloc(reinterpret_cast<void**>(&s->fm_host), m * sizeof(double)) != hipSuccess) {
        s->fm_host = nullptr;
        return set_err(s, MSM_ERR_NOMEM, "pinned staging allocation failed");
    }
    s->fm_cap_host = m;
    return MSM_OK;
}

// The spline model's shape checks: one pair kind (one embedding kind for EAM) and ld covering
// both grids.
static int fm_check_model(msm_sim* s, int ld) {
    if (s->tab.nkind != 1 || (s->tab.pot_type != 0 && s->tab.nkind1 != 1))
        return set_err(s, MSM_ERR_ARG, "the spline model covers tables of one kind");
    const int m = std::max(s->tab.ntab2 - 2, s->tab.pot_type != 0 ? s->tab.nembd : 0);
    if (ld < m) return set_err(s, MSM_ERR_ARG, "ld smaller than NTAB or NEMBD");
    return MSM_OK;
}

int msm_set_spline_table(msm_sim* s, const double* coef, int ld, int coef_on_device) {
    if (!s || !coef) return MSM_ERR_ARG;
    if (!s->has_table) return set_err(s, MSM_ERR_STATE, "no force table (msm_set_force_table)");
    int rc = fm_check_model(s, ld);
    if (rc != MSM_OK) return rc;
    (void)hipSetDevice(s->device);
    if (!s->fm_ev) HIPCHK(s, hipEventCreateWithFlags(&s->fm_ev, hipEventDisableTiming));
    // the staging buffer and the table buffers are free once the previous build has run
    if (s->fm_pending) HIPCHK(s, hipEventSynchronize(s->fm_ev));
    s->fm_pending = false;
    const size_t m = 3 * static_cast<size_t>(ld);
    const double* c = coef;
    if (!coef_on_device) {
        if ((rc = fm_reserve(s, &s->fm_coef, m, s->fm_cap_coef)) != MSM_OK) return rc;
        if ((rc = fm_reserve_host(s, m + 1)) != MSM_OK) return rc;
        std::memcpy(s->fm_host, coef, m * sizeof(double));
        HIPCHK(s, hipMemcpyAsync(s->fm_coef, s->fm_host, m * sizeof(double),
                                 hipMemcpyHostToDevice, s->stream));
        c = s->fm_coef;
    }
    // same buffers, sizes and grid: the table descriptor and captured segments stay valid
    msm::launch_spline_tables(c, ld, s->tab, s->d_b2, s->d_f4, s->d_e4, s->d_m4, s->stream);
    HIPCHK(s, hipGetLastError());
    // stream-ordered before every later call on this context; msm_set_force_table, whose
    // copies are not, and the next build wait for the event.  A caller's device coefficients
    // are read before the call returns.
    if (coef_on_device) {
        HIPCHK(s, hipStreamSynchronize(s->stream));
        return MSM_OK;
    }
    HIPCHK(s, hipEventRecord(s->fm_ev, s->stream));
    s->fm_pending = true;
    return MSM_OK;
}

int msm_fm_set_targets(msm_sim* s, const double* f_target, const double* e_target, double w_e) {
    if (!s || !f_target || !e_target) return MSM_ERR_ARG;
    if (!s->has_atoms || !s->has_boxes) return set_err(s, MSM_ERR_STATE, "no atoms or boxes");
    if (!(w_e >= 0.0)) return set_err(s, MSM_ERR_ARG, "w_e must be non-negative");
    (void)hipSetDevice(s->device);
    const size_t n3 = 3 * static_cast<size_t>(std::max(s->n, 1));
    const size_t nbox = s->boxes.size();
    int rc = fm_reserve(s, &s->fm_ft, n3, s->fm_cap_ft);
    if (rc == MSM_OK) rc = fm_reserve(s, &s->fm_et, nbox, s->fm_cap_et);
    if (rc != MSM_OK) return rc;
    HIPCHK(s, hipMemcpyAsync(s->fm_ft, f_target, 3 * static_cast<size_t>(s->n) * sizeof(double),
                             hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipMemcpyAsync(s->fm_et, e_target, nbox * sizeof(double), hipMemcpyHostToDevice,
                             s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->fm_n = s->n;
    s->fm_w_e = w_e;
    return MSM_OK;
}

int msm_fm_loss_grad(msm_sim* s, double* loss, double* coef_grad, int ld, int grad_on_device) {
    int rc = check_ready(s, true);
    if (rc != MSM_OK) return rc;
    if (!loss) return set_err(s, MSM_ERR_ARG, "null argument");
    if (s->fm_n != s->n || s->n < 1)
        return set_err(s, MSM_ERR_STATE, "no targets for these atoms (msm_fm_set_targets)");
    if (coef_grad && (rc = fm_check_model(s, ld)) != MSM_OK) return rc;
    (void)hipSetDevice(s->device);
    const int n = s->n, nbox = static_cast<int>(s->boxes.size());
    size_t nb1 = 0, nb2 = 0, stride = 0;
    if ((rc = adj_reserve(s, nb1, nb2, stride)) != MSM_OK) return rc;
    if ((rc = fm_reserve(s, &s->fm_part, msm::fm_parts(n), s->fm_cap_part)) != MSM_OK) return rc;
    if ((rc = fm_reserve(s, &s->fm_loss, 1, s->fm_cap_loss)) != MSM_OK) return rc;
    const size_t m = coef_grad && !grad_on_device ? 3 * static_cast<size_t>(ld) : 0;
    if (m && (rc = fm_reserve(s, &s->fm_grad, m, s->fm_cap_grad)) != MSM_OK) return rc;
    if ((rc = fm_reserve_host(s, m + 1)) != MSM_OK) return rc;
    // everything up to the flags as one graph on repeated calls (one launch instead of about
    // twenty); the key holds what the captured kernels and copies bake in
    hipError_t ce = hipSuccess;
    auto enqueue = [&]() {
        clear_flags(s);
        enqueue_density(s);
        enqueue_force(s, 0, 1, 0.0);
        AtomsDev a = atoms_view(s, s->scur, s->pcur);
        // with a gradient the energy pass also keeps rho for the adjoint, whose F' is the
        // density pass's (in pos.w), so the adjoint's own rho pass is skipped
        msm::launch_epot(a, list_view(s), s->tab, s->boxes[0], s->d_boxes, multibox(s), s->d_err,
                         s->stream, coef_grad ? s->adj_rho : nullptr);
        msm::launch_box_sum(s->epot, s->sb.start, s->d_boxes, nbox, s->d_ebox, s->stream);
        const AtomSet& z = s->set[s->scur];
        const size_t na = s->n_alloc;
        msm::launch_fm_loss(z.f, z.f + na, z.f + 2 * na, z.gid, n, s->fm_ft, s->d_ebox, s->fm_et,
                            s->d_boxes, nbox, s->fm_w_e, s->adj_dldf, s->adj_gbox, s->fm_part,
                            s->fm_loss, s->stream);
        if (coef_grad) {
            adj_launch(s, nb1, stride, true);
            double* g = grad_on_device ? coef_grad : s->fm_grad;
            msm::launch_spline_grad(s->adj_out, s->adj_out + nb1, s->tab, ld, g, s->stream);
            if (m && ce == hipSuccess)
                ce = hipMemcpyAsync(s->fm_host + 1, g, m * sizeof(double), hipMemcpyDeviceToHost,
                                    s->stream);
        }
        if (ce == hipSuccess)
            ce = hipMemcpyAsync(s->fm_host, s->fm_loss, sizeof(double), hipMemcpyDeviceToHost,
                                s->stream);
    };
    if (s->timing) {  // direct launches, each bracketed by events
        enqueue();
        HIPCHK(s, ce);
    } else {
        auto P = [](const void* q) { return static_cast<uint64_t>(reinterpret_cast<uintptr_t>(q)); };
        uint64_t webits = 0;
        std::memcpy(&webits, &s->fm_w_e, sizeof(webits));
        const std::vector<uint64_t> key = {
            s->graph_gen, static_cast<uint64_t>(s->scur), static_cast<uint64_t>(s->pcur),
            static_cast<uint64_t>(s->track_dis), static_cast<uint64_t>(n),
            static_cast<uint64_t>(ld), coef_grad ? 1u : 0u, static_cast<uint64_t>(m),
            grad_on_device ? P(coef_grad) : 0u, webits, P(s->adj_dldf), P(s->adj_gbox),
            P(s->adj_g), P(s->adj_rho), P(s->adj_lam), P(s->adj_part), P(s->adj_out),
            P(s->fm_part), P(s->fm_loss), P(s->fm_grad), P(s->fm_host), P(s->fm_ft),
            P(s->fm_et), P(s->epot), P(s->d_ebox)};
        auto it = s->fm_graphs.find(key);
        if (it == s->fm_graphs.end()) {
            if (s->fm_graphs.size() >= 4) {  // buffers or sizes moved on: drop the old ones
                for (auto& kv : s->fm_graphs) (void)hipGraphExecDestroy(kv.second);
                s->fm_graphs.clear();
            }
            HIPCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            enqueue();
            hipGraph_t gr = nullptr;
            const hipError_t ee = hipStreamEndCapture(s->stream, &gr);
            if (ce == hipSuccess) ce = ee;
            hipGraphExec_t ge = nullptr;
            if (ce == hipSuccess) ce = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
            if (gr) (void)hipGraphDestroy(gr);
            if (ce != hipSuccess)
                return set_err(s, MSM_ERR_HIP, std::string("fitting-loop graph: ") +
                                                   hipGetErrorString(ce));
            it = s->fm_graphs.emplace(key, ge).first;
        }
        HIPCHK(s, hipGraphLaunch(it->second, s->stream));
    }
    rc = read_flags(s);  // synchronises the stream
    *loss = s->fm_host[0];
    if (m) std::memcpy(coef_grad, s->fm_host + 1, m * sizeof(double));
    return rc;
}

int msm_list_stats(msm_sim* s, long long* total, int* maxk) {
    if (!s || !total || !maxk) return MSM_ERR_ARG;
    if (!s->has_list) return set_err(s, MSM_ERR_STATE, "no neighbour list");
    (void)hipSetDevice(s->device);
    msm::launch_fill_int(reinterpret_cast<int*>(s->d_stats), 0, 4, s->stream);
    msm::launch_list_stats(s->kvois, s->n, s->d_stats,
                           reinterpret_cast<int*>(s->d_stats + 1), s->stream);
    unsigned long long h[2] = {0, 0};
    HIPCHK(s, hipMemcpyAsync(h, s->d_stats, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    *total = static_cast<long long>(h[0]);
    *maxk = static_cast<int>(h[1] & 0xffffffffu);
    return MSM_OK;
}

}  // extern "C"
