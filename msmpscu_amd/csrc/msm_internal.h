// msm_internal.h -- device data layout and kernel launchers shared by the HIP translation units
// of libmsm_hip.so.
//
// HBM layout (all per-atom arrays in cell-sorted order, rebuilt every neighbour update):
//   pos[2][n]   double4 {x, y, z, w}; w holds F'(rho_i) after the density pass, so the force pass
//               gathers a neighbour's position and F' with one 32-byte load.  Two buffers: the
//               fused force+integrator kernel reads one and writes the next positions to the other.
//   v, f, dis   three SoA double arrays each (x|y|z, stride n_alloc), owned by one thread per atom.
//   ityp, statu, boxid, gid, cell   int32 per atom (gid = caller's index of the sorted atom).
//   kvois[n], indi[mx][n]           full neighbour list, slot-major (coalesced per slot).
//   tables      kind-major knots 0..ntab+1 holding {value, next-minus-value} pairs, so one
//               lerp v(KK) + DK*(v(KK+1)-v(KK)) is one 16-byte load and one multiply-add.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace msm {

constexpr int kMaxGroup = 16;
constexpr int kBlock = 256;

// kErrDt: flag of the time-step check (not an error; read by msm_check_timestep alone)
enum ErrSlot { kErrOverflow = 0, kErrOutOfBox = 1, kErrEmbd = 2, kErrDt = 3, kErrSlots = 4 };

struct BoxDev {
    double lo[3], len[3], hlen[3];
    double ilen[3];  // 1/len for periodic directions, 0 otherwise (branch-free minimum image)
    int ifpd[3];
    int ncell[3];
    int cell_off;  // first global cell id of this box
    int natoms;    // atoms assigned to this box
    int bdim[3];   // cells per brick of the LDS-staged list build
    int nbrick[3]; // bricks per direction
    int brick_off; // first global brick id of this box
    int cbs;       // cell ordering: bricks of cbs^3 cells are contiguous (1 = row-major)
};

// LDS-staged list build: threads per workgroup, most halo atoms per workgroup (the density and
// force brick kernels stage at most kNlCapPass of them) and halo cells per brick (bricks <= 4^3).
constexpr int kNlBlock = 1024;
constexpr int kNlCap = 4096;
constexpr int kNlCapPass = 1536;
constexpr int kNlMaxHalo = 216;

struct TabDev {
    const double2* b2;  // density table {POTB, dPOTB}, nkind x (ntab+2)
    const double4* f4;  // force table {FPOTR, dFPOTR, FPOTB, dFPOTB}
    const double4* e4;  // energy table {POTR, dPOTR, POTB, dPOTB}
    const double4* m4;  // embedding {FEMBD, dFEMBD, DFEMBD, dDFEMBD}, nkind1 x (nembd+2)
    int ntab2, nembd2, nembd, ngroup;
    int nkind;          // table kinds (rows of b2/f4/e4)
    int nkind1;         // embedding kinds (rows of m4; 1 for FS tables)
    double csi, rhod, ru2max;
    int pot_type;
    int kpair[kMaxGroup * kMaxGroup];  // 0-based kinds, [ti*ngroup + tj]
    int kembd[kMaxGroup];
    double cm[kMaxGroup];
};

struct AtomsDev {
    double4* pos;      // positions read by this launch (w = F')
    double4* pos_out;  // next positions (fused integrator)
    double *vx, *vy, *vz;
    double *fx, *fy, *fz;
    double *dx, *dy, *dz;
    int* ityp;
    int* statu;
    int* boxid;
    double* epot;
    int n;
};

struct ListDev {
    const int* kvois;
    const int* indi;            // neighbour sorted index, slot-major
    const unsigned short* sidx; // neighbour LDS slot in the home brick's halo (brick kernels)
    int* cidx;                  // in-cutoff pairs of the current step (global-list path), or null
    int* ckv;                   // their count per lane (n_alloc * lanes per atom)
    unsigned long long* cmask;  // LDS-table passes: per lane (atom * lanes + lane), bit k set if
                                // the lane's k-th slot is inside the cutoff this step (written
                                // by the density pass for the force pass), or null
    int stride;                 // distance between consecutive slots of one atom (kListRow)
    int mx;                     // slots per atom (MXKVOIS)
    int scap;                   // brick passes stage a brick (and read sidx) only if its halo
                                // holds at most this many atoms: the list build wrote sidx for
                                // those bricks (0 for a list from the global scan)
};

// Neighbour lists are stored in blocks of kListRow atoms: slot iw of atom p lives at
// (p / kListRow) * kListRow * MXKVOIS + iw * kListRow + p % kListRow, so the slot loads of a wave
// are contiguous and all slots of a wave's atoms share one small address range.
constexpr int kListRow = 64;

struct StepParams {
    double h, hs2, h2s2;
    int mode;      // 0: force only, 1: force + corrector, 2: force + corrector + next predictor
    int write_f;   // store forces
    int track_dis; // accumulate DIS in the fused predictor
};

struct SortBufs {
    int* key;    // cell id per atom (current order)
    int* slot;   // arrival slot within the cell
    int* count;  // ncell_total + 2
    int* start;  // ncell_total + 2 (exclusive scan of count)
    int* perm;   // sorted position -> current index
    int* scan_tmp;
};

// ---- launchers (msm_kernels.hip) ----
void launch_fill_int(int* p, int v, int n, hipStream_t s);
void launch_cell_assign(const double4* pos, const int* statu, const int* boxid,
                        const BoxDev* boxes, int n, int ncell_total, SortBufs sb, int* err,
                        hipStream_t s);
void launch_scan_exclusive(const int* in, int* out, int n, int* tmp, hipStream_t s);
void launch_scatter_sort(int n, int ncell_total, SortBufs sb, const int* gid, hipStream_t s);
void launch_gather(int n, int n_alloc, const int* perm, const AtomsDev& src, const int* gid_src,
                   AtomsDev& dst, int* gid_dst, const int* key, int* cell_dst, int keep_f,
                   hipStream_t s);
void launch_nlist(int n, const double4* pos, const int* ityp, const int* statu, const int* boxid,
                  const int* cell, const int* start, const BoxDev* boxes, int ngroup,
                  const double* rnl2, int mx, int* kvois, int* indi, int stride, int* err,
                  hipStream_t s);
// LDS-staged variants over bricks of cells (msm_bricks.hip); out-of-box atoms (sentinel bucket)
// get KVOIS = 0, F' = 0, F = 0 and keep their position.
size_t nlist_lds_bytes(int cap, int ngroup);  // dynamic LDS of the list build for cap halo atoms
void launch_nlist_bricks(int wide, int nbricks, int lanes, int cap, int write_sidx,
                         const int* brick_box, int n,
                         int ncell_total,
                         const double4* pos, const int* ityp, const int* start,
                         const BoxDev* boxes, int ngroup, const double* rnl2, int mx, int* kvois,
                         int* indi, unsigned short* sidx, int stride, int* err, hipStream_t s);
void launch_density_bricks(int nbricks, const int* brick_box, const AtomsDev& a,
                           const ListDev& l, const TabDev& t, const BoxDev* boxes,
                           const int* start, int ncell_total, int* err, hipStream_t s);
void launch_force_bricks(int nbricks, const int* brick_box, const AtomsDev& a, const ListDev& l,
                         const TabDev& t, const BoxDev* boxes, const int* start, int ncell_total,
                         const StepParams& sp, hipStream_t s);
// Atoms of the out-of-box bucket (from *first to n): F' = 0 (clear_den), F = 0 (write_f), position
// carried to pos_out (copy_pos).
void launch_outside_atoms(const AtomsDev& a, const int* first, int write_f, int copy_pos,
                          int clear_den, hipStream_t s);
// Global-list passes with tpa (1, 2, 4 or 8) lanes per atom (msm_tpa.hip).
void launch_density_tpa(int tpa, const AtomsDev& a, const ListDev& l, const TabDev& t,
                        const BoxDev& box0, const BoxDev* boxes, int multibox, int* err,
                        hipStream_t s);
void launch_force_tpa(int tpa, const AtomsDev& a, const ListDev& l, const TabDev& t,
                      const BoxDev& box0, const BoxDev* boxes, int multibox, const StepParams& sp,
                      hipStream_t s);
// Global-list passes with the single table kind's knots in LDS (msm_ldstab.hip): a window of
// the top wf (force, {FPOTR, FPOTB}) / wd (density, POTB) knots starting at k0f / k0d, as large
// as the LDS share of one workgroup allows; persistent grids of block_* threads and wgs_*
// workgroups per CU.
constexpr int kLtMaxBlock = 1024;
struct LtWin {
    int k0f, wf, k0d, wd;
    int sweep;  // chunk order of the persistent grid (LtShape::sweep)
};
struct LtShape {
    int block_d, wgs_d, block_f, wgs_f;
    // 0: each workgroup walks a contiguous range of its XCD's chunks; 1: the workgroups of an
    // XCD take its chunks round-robin, so the chunks in flight on one XCD are adjacent
    int sweep;
};
LtWin lt_window(const TabDev& t, int wgs_d, int wgs_f);
void launch_density_lt(int tpa, const LtShape& sh, const AtomsDev& a, const ListDev& l,
                       const TabDev& t, const BoxDev& box0, const BoxDev* boxes, int multibox,
                       int* err, hipStream_t s);
void launch_force_lt(int tpa, const LtShape& sh, const AtomsDev& a, const ListDev& l,
                     const TabDev& t, const BoxDev& box0, const BoxDev* boxes, int multibox,
                     const StepParams& sp, hipStream_t s);
// LDS bricks with the table window in LDS (msm_ldstab.hip): tpa (1, 2 or 4) lanes per home
// atom, persistent grids over the bricks of the list build (whose slot list they read); the
// staged halo capacity (at most the list build's, l.scap) and the windows share the LDS.
struct BtPlan {
    LtWin w;
    int cap_d, cap_f;  // halo atoms a brick may have to be staged (density, force pass)
};
BtPlan bt_plan(const TabDev& t, const LtShape& sh, int scap);
void launch_density_bt(int tpa, const LtShape& sh, int nbricks, const int* brick_box,
                       const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev* boxes,
                       const int* start, int ncell_total, int* err, hipStream_t s);
void launch_force_bt(int tpa, const LtShape& sh, int nbricks, const int* brick_box,
                     const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev* boxes,
                     const int* start, int ncell_total, const StepParams& sp, hipStream_t s);
// Staged-brick passes (msm_stage.hip): one brick's halo staged in LDS per workgroup step, the
// pairs read as halo slots; halo capacity cap (the list build's) and knot windows in the rest of
// the LDS share.  ok: both windows hold at least a quarter of the table.
struct SbPlan {
    LtWin w;
    int cap;
    bool ok;
};
// Slot loads the staged passes may issue past an atom's last list slot (allocation slack of
// the slot list, in entries).
constexpr int kSbSlack = 32 * kListRow;
SbPlan sb_plan(const TabDev& t, const LtShape& sh, int scap);
void launch_density_sb(const LtShape& sh, int nbricks, const int* brick_box, const AtomsDev& a,
                       const ListDev& l, const TabDev& t, const BoxDev* boxes, const int* start,
                       int ncell_total, int* err, hipStream_t s);
void launch_force_sb(const LtShape& sh, int nbricks, const int* brick_box, const AtomsDev& a,
                     const ListDev& l, const TabDev& t, const BoxDev* boxes, const int* start,
                     int ncell_total, const StepParams& sp, hipStream_t s);
// Reverse-mode table gradients (msm_adjoint.hip): scratch of the adjoint passes.
struct AdjDev {
    double* g;           // dL/dF in the sorted order, (x, y, z, 0) per atom, 32-byte aligned
                         // (zero for inactive atoms)
    const double* gbox;  // dL/dE per box
    double* rho;         // rho_k
    double* lam;         // dL/dF'_k
    double* part;        // per-workgroup bins, part_stride doubles each
    size_t part_stride;
};
int adjoint_workgroups(int n);
// Bin counts of the two adjoint passes (doubles): pairs nb1, density nb2.
void adjoint_bins(const TabDev& t, int& nb1, int& nb2);
// out_pairs: 3 x nkind x ntab2 knot gradients [FPOTR | FPOTB | POTR], each kind-major; out_dens:
// nkind x ntab2 + 2 x nkind1 x nembd2 [POTB | FEMBD | DFEMBD] (embedding parts for EAM only).
// rho_ready: d.rho and pos.w (F') already hold the current values (the energy pass's rho_out and
// the density pass), so the rho pass is skipped.
void launch_adjoint(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                    const BoxDev* boxes, int multibox, const double* dldf, const int* gid,
                    AdjDev d, double* out_pairs, double* out_dens, hipStream_t s,
                    bool rho_ready = false);
// Device-resident force matching (msm_adjoint.hip).  launch_spline_tables: the knot tables b2,
// f4, e4 (and m4 for EAM) of one kind from the cubic B-spline coefficients coef (rows pair,
// density, embedding of ld doubles) on t's grid.  launch_fm_loss: from the sorted forces and the
// box energies, dL/dF (caller order, FP(N,3)) and dL/dE for the adjoint and the loss into *loss
// (part: fm_parts(n) doubles of scratch).  launch_spline_grad: dL/dc (3 rows of ld) from the
// adjoint's bins of one kind.
void launch_spline_tables(const double* coef, int ld, const TabDev& t, double2* b2, double4* f4,
                          double4* e4, double4* m4, hipStream_t s);
int fm_parts(int n);
void launch_fm_loss(const double* fx, const double* fy, const double* fz, const int* gid, int n,
                    const double* ft, const double* ebox, const double* et, const BoxDev* boxes,
                    int nbox, double w_e, double* dldf, double* gbox, double* part, double* loss,
                    hipStream_t s);
void launch_spline_grad(const double* pairs, const double* dens, const TabDev& t, int ld,
                        double* g, hipStream_t s);
// Two lanes per atom with 32-byte records gathered by lane pairs (msm_pair.hip); the compacted
// list is not used there.
void launch_density_pair(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                         const BoxDev* boxes, int multibox, int* err, hipStream_t s);
void launch_force_pair(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                       const BoxDev* boxes, int multibox, const StepParams& sp, hipStream_t s);
void launch_density(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                    const BoxDev* boxes, int multibox, int* err, hipStream_t s);
void launch_force(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                  const BoxDev* boxes, int multibox, const StepParams& sp, hipStream_t s);
// rho_out (may be null): each atom's density sum rho (0 for inactive atoms) as well.
void launch_epot(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                 const BoxDev* boxes, int multibox, int* err, hipStream_t s,
                 double* rho_out = nullptr);
void launch_predictor(const AtomsDev& a, const TabDev& t, const BoxDev& box0,
                      const BoxDev* boxes, int multibox, double h, hipStream_t s);
void launch_correction(const AtomsDev& a, const TabDev& t, double h, hipStream_t s);
void launch_unsort_d(const double* src, double* dst, const int* gid, int n, hipStream_t s);
void launch_sort_pos(const double* src, double4* dst, const int* gid, int n, hipStream_t s);
void launch_sort_d(const double* src, double* dst, const int* gid, int n, hipStream_t s);
// dst[p] += src[gid[p]] (a caller-order array added into the sorted order)
void launch_add_sorted(const double* src, double* dst, const int* gid, int n, hipStream_t s);
void launch_unsort_i(const int* src, int* dst, const int* gid, int n, hipStream_t s);
void launch_unsort_pos(const double4* src, double* dst, const int* gid, int n, int comp,
                       hipStream_t s);
void launch_unsort_list(const int* kvois, const int* indi, int stride, const int* gid, int n,
                        int mx, int* kv_out, int* indi_out, int ld, hipStream_t s);
// Virial tensor per atom (6 components, component-major, stride a.n) and forces.
void launch_virial(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                   const BoxDev* boxes, int multibox, double* vir, hipStream_t s);
// Per-box virial tensors without a per-atom buffer: nchunk >= ceil(atoms in any box's cells /
// kBlock); part holds nbox * nchunk * 6 partial sums, out the 9 * nbox tensor entries
// (column-major per box, as msm_calc_ptensor returns them).  Forces are written as by
// launch_virial.
void launch_virial_boxes(const AtomsDev& a, const ListDev& l, const TabDev& t,
                         const BoxDev& box0, const BoxDev* boxes, int multibox, const int* start,
                         int nbox, int ncell_total, int nchunk, double* part, double* out,
                         hipStream_t s);
// Thermalisation and quench (MD_DiffScheme_GPU.F90): DAMPING, per-atom EK (-1e32 for atoms that
// do not count), per-box sum and count of non-negative EK, velocity scaling by the factor of the
// atom's box, and max |epot - epot0| (then epot0 = epot; *out zero on entry, a double's bits).
void launch_damping(const AtomsDev& a, hipStream_t s);
void launch_check_dt(const AtomsDev& a, const TabDev& t, double th, double mxd2, int* flag,
                     hipStream_t s);
void launch_ekin(const AtomsDev& a, const TabDev& t, double* ek, hipStream_t s);
void launch_box_ekin(const double* ek, const int* start, const BoxDev* boxes, int nbox,
                     double* sum, double* cnt, hipStream_t s);
void launch_velscale(const AtomsDev& a, const double* scal, hipStream_t s);
void launch_epot_maxdelta(const double* epot, double* epot0, int n, unsigned long long* out,
                          hipStream_t s);

// Thermalizing_MC_DEV (msm_thermal.hip): Maxwell velocities from Philox4x32-10 keyed by seed,
// counter (caller index gid, component, draw), then per-box removal of the mass-weighted mean
// velocity; mom holds 4 doubles per box (sum m v, M).
void launch_thermalize(const AtomsDev& a, const TabDev& t, const int* gid, double tkb,
                       unsigned long long seed, unsigned long long draw, const int* start,
                       const BoxDev* boxes, int nbox, int ncell_total, double* mom,
                       hipStream_t s);

// steepest-descent quench helpers (3-component vectors as three arrays of n doubles)
struct P3 {
    double* x;
    double* y;
    double* z;
};
constexpr int kQuenchBlocks = 1024;   // most workgroups of the dot-product reductions
int quench_blocks(int n);             // workgroups (and partials per component) used for n
void launch_dot3(P3 u, P3 v, int n, double* part, double* out, hipStream_t s);
void launch_dxdf(P3 dx, P3 p, P3 f, int n, double* part, double* out, hipStream_t s);
void launch_scale3(P3 src, double c, int n, P3 dst, hipStream_t s);
void launch_dir_update(P3 f, double g, int n, P3 d, hipStream_t s);
void launch_maxabs3(P3 v, int n, unsigned long long* out, hipStream_t s);
void launch_move3(const AtomsDev& a, P3 dx, const BoxDev& box0, const BoxDev* boxes, bool mb,
                  hipStream_t s);
// AVP(N, 9) in caller order (column-major) from launch_virial's per-atom halves.
void launch_unsort_stress(const double* vir, double* avp, const int* gid, int n, hipStream_t s);
void launch_box_sum(const double* epot, const int* start, const BoxDev* boxes, int nbox,
                    double* out, hipStream_t s);
void launch_list_stats(const int* kvois, int n, unsigned long long* total, int* maxk,
                       hipStream_t s);

}  // namespace msm
