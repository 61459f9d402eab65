/*
 * orc_md.h -- CPU oracle for the tabulated FS/EAM molecular-dynamics force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker against which the
 * HIP implementation is compared.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never links it.
 *
 * It is a plain-C restatement (FP64, no FMA contraction) of the reference's
 * host-executed routines:
 *   table build      MDLIB/sor/Common/MD_TypeDef_ForceTable.F90:530-622, 890-1056, 1151-1230
 *   W FS potential   MDLIB/sor/Potentials/EM_TB_WangJun_W-HE_2010/FS_Ackland_WW.F90:23-91
 *   FS embedding     MDLIB/sor/Common/MD_TypeDef_ForceTable.F90:1859-1882
 *   neighbour lists  MDLIB/sor/Common/MD_NeighborsList.F90:52-148, 152-391, 396-633
 *   FS force (half)  MDLIB/sor/Common/MD_FS_ForceTable.F90:150-395, 646-669
 *   full-list force  MDLIB/sor/CommonGPU/MD_FS_ForceTable_GPU.F90:341-516, 623-806, 1451-1609
 *                    MDLIB/sor/CommonGPU/MD_EAM_ForceTable_GPU.F90:369-548, 646-829, 1470-1636
 *   integrator       MDLIB/sor/Common/MD_SwopeScheme.F90:22-283
 *   step order       MDLIB/sor/Appshell/MD_Method_GenericMD_GPU.F90:596-627
 *   RNG / Maxwell    LIB/sor/f/RandGenerators/DRAND32.F90:60-85,
 *                    MDLIB/sor/Common/MD_TypeDef_SimBox.F90:4865-4942
 *
 * Conventions (shared with the product's C-ABI):
 *   - positions/velocities/forces are SoA "XP(N,3)" column-major: x at [i], y at [i+n], z at [i+2n];
 *   - atom groups (types) are 0-based;
 *   - neighbour lists are slot-major INDI(N,MXKVOIS): neighbour iw of atom i at [i + iw*n], 0-based ids;
 *   - force tables are kind-major with one zero pad entry on each side:
 *     T(k,i) for i = 0..ntab+1 at [k*(ntab+2) + i]; T(k,0) = T(k,ntab+1) = 0;
 *   - KPAIR(ti,tj) at kpair[ti*ngroup + tj] (0-based kinds); KEMBD(t) at kembd[t];
 *   - the box shape matrix is the identity (orthorhombic boxes), for which the reference's
 *     BOXSHAPE products are exact no-ops.
 */
#ifndef ORC_MD_H
#define ORC_MD_H

#ifdef __cplusplus
extern "C" {
#endif

/* status bits, MDLIB/sor/Common/MD_Const.F90:79-99 */
#define ORC_STATU_ACTIVE    1
#define ORC_STATU_FIXPOSX   2
#define ORC_STATU_FIXPOSY   4
#define ORC_STATU_FIXPOSZ   8
#define ORC_STATU_FIXVELX   16
#define ORC_STATU_FIXVELY   32
#define ORC_STATU_FIXVELZ   64
#define ORC_STATU_OUTOFBOX  65536
#define ORC_STATU_REFLECT   131072
#define ORC_STATU_TRANSMIT  262144
#define ORC_STATU_PASSBOUND 524288

#define ORC_OK             0
#define ORC_ERR_OVERFLOW   1   /* more neighbours than mxkvois */
#define ORC_ERR_OUTOFBOX   2   /* active atom outside the cell grid */
#define ORC_ERR_ARG        3
#define ORC_ERR_NOMEM      4

typedef void (*orc_pairfn)(const double* r, double* v, double* dv);

/* ---------------- potentials (eV, Angstrom) ---------------- */
/* W Finnis-Sinclair with the polynomial pair term only (coefficients as given for the
 * benchmark: c=3.25, 47.1346, -33.7666, 6.2542; rho = A^2 (r-4.4002)^2, A = 1.8964).
 * NN returns (0.5 V, -dV/dr); NE returns (rho, -drho/dr). */
void orc_pot_wfs_nn(const double* r, double* potr, double* fpotr);
void orc_pot_wfs_ne(const double* r, double* potb, double* fpotb);
/* Ackland-Thetford W-W with the reference's coefficients including the core term
 * (FS_Ackland_WW.F90:11-91), converted to eV/Angstrom. */
void orc_pot_ackland_ww_nn(const double* r, double* potr, double* fpotr);
void orc_pot_ackland_ww_ne(const double* r, double* potb, double* fpotb);
/* FS embedding F = -sqrt(rho), F' = -1/(2 sqrt(rho)), 0 at rho <= 0 (MD_TypeDef_ForceTable.F90:1859-1882) */
void orc_emb_fs(const double* rho, double* f, double* df);

/* ---------------- tables ---------------- */
/* Create_Pairwise_ForceTable for one kind.  Writes entries 1..ntab of the padded row
 * (row must hold ntab+2 doubles; pads are zeroed).  rhomx_inout is the running maximum of
 * POTB*rhomult (starts at -1 in the reference). ne may be NULL. */
void orc_table_pairwise(int ntab, double rmax, orc_pairfn nn, orc_pairfn ne, double rhomult,
                        double* rhomx_inout, double* potr, double* fpotr, double* potb, double* fpotb);
/* Create_EMBDFUNTable for one kind: F(k,i) at rho=(i-1)*rhod, i=1..nembd (padded row). */
void orc_table_embd(int nembd, double rhod, orc_pairfn embd, double* fembd, double* dfembd);

typedef struct orc_ftable {
    int pot_type;            /* 0 = FS_TYPE, 1 = EAM_TYPE */
    int ngroup, nkind, ntab;
    double csi;
    const int* kpair;
    const double *potr, *fpotr, *potb, *fpotb;
    int nkind1, nembd;
    double rhod;
    const int* kembd;
    const double *fembd, *dfembd;
} orc_ftable;

/* ---------------- neighbour lists ---------------- */
/* Cal_NeighboreListC (half, 14-cell stencil).  Returns ORC_OK or an error code. */
int orc_nlist_half(int n, const double* xp, const int* ityp, const int* statu,
                   const double lo[3], const double len[3], const int ifpd[3],
                   int ngroup, const double* nb_rm, int mxkvois, int* kvois, int* indi);
/* Cal_NeighboreList2C (full, 27-cell stencil). */
int orc_nlist_full(int n, const double* xp, const int* ityp, const int* statu,
                   const double lo[3], const double len[3], const int ifpd[3],
                   int ngroup, const double* nb_rm, int mxkvois, int* kvois, int* indi);
/* Brute-force full list with the minimum-image separation (Cal_NeighboreList predicate,
 * indexed by type pair), used to pin the cell lists. */
int orc_nlist_brute(int n, const double* xp, const int* ityp, const int* statu,
                    const double len[3], const int ifpd[3],
                    int ngroup, const double* nb_rm, int mxkvois, int* kvois, int* indi);

/* ---------------- forces ---------------- */
/* CALFORCE_FS_Force_Table: half list, Newton's third law, per-pair cutoff RU(ti,tj).
 * Outputs FP(N,3), EPOT(N) = ER - sqrt(DEN), DEN(N) (density), DDEN(N) (F'), VTENSOR(3,3) and,
 * when vabs is not NULL, the sum of the absolute values of VTENSOR's pair terms per component
 * (the scale of its rounding error). */
int orc_force_fs_half(int n, const double* xp, const int* ityp, const int* kvois, const int* indi,
                      int mxkvois, const double len[3], const int ifpd[3], const double* ru,
                      const orc_ftable* t, double* fp, double* epot, double* den, double* dden,
                      double* vtensor, double* vabs);
/* CAL_FS/EAM_AtomicStress_KERNEL: per-atom stress AP(N,9) column-major (P11, P12, ..., P33), the
 * full-list sum of SEP_a SEP_b FORTOT for active atoms (zero otherwise) given F' in dden (as
 * orc_force_full returns it); apabs (may be NULL) gets the sums of the terms' absolute values. */
int orc_atomic_stress_full(int n, const double* xp, const int* ityp, const int* statu,
                           const int* kvois, const int* indi, const double len[3],
                           const int ifpd[3], double ru2max, const orc_ftable* t,
                           const double* dden, double* ap, double* apabs);
/* Full-list restatement of the device kernels PRECALFOR/CALFORCE/CALEPOT (FS or EAM per
 * t->pot_type): single cutoff max(RU^2), only ACTIVE atoms compute, KPAIR/KPAIR^T pairing as
 * in the corrected kernels.  Outputs FP, DEN (F'), EPOT. */
int orc_force_full(int n, const double* xp, const int* ityp, const int* statu, const int* kvois,
                   const int* indi, int mxkvois, const double len[3], const int ifpd[3],
                   double ru2max, const orc_ftable* t, double* fp, double* dden, double* epot);

/* Reverse mode of orc_force_full: given G = dL/dFP (N,3) and g = dL/dE (E = sum of EPOT over the
 * atoms), accumulates dL/d(knot) into the padded kind-major rows g_potr, g_fpotr, g_potb,
 * g_fpotb (nkind x (ntab+2)) and, for EAM, g_fembd, g_dfembd (nkind1 x (nembd+2); may be NULL
 * for FS), every kind that KPAIR and KEMBD select.  Outputs are added to, not overwritten. */
int orc_adjoint_full(int n, const double* xp, const int* ityp, const int* statu,
                     const int* kvois, const int* indi, const double len[3], const int ifpd[3],
                     double ru2max, const orc_ftable* t, const double* dldf, double dlde,
                     double* g_potr, double* g_fpotr, double* g_potb, double* g_fpotb,
                     double* g_fembd, double* g_dfembd);

/* ---------------- integrator ---------------- */
void orc_predictor(int n, double* xp, double* xp1, const double* fp, double* dis, int* statu,
                   const int* ityp, const double* cm, double h, const double lo[3],
                   const double len[3], const int ifpd[3]);
/* DAMPING_KERNEL, CALEKIN_KERNEL and VelScaling (MD_DiffScheme_GPU.F90:125-187, 851-904,
 * 1262-1448) for one box; orc_velscale returns ORC_ERR_ARG when the box has no kinetic energy. */
void orc_damping(int n, double* xp1, const double* fp, const int* statu);
void orc_ekin(int n, const double* xp1, const int* statu, const int* ityp, const double* cm,
              double* ek);
int orc_velscale(int n, double* xp1, const int* statu, const int* ityp, const double* cm,
                 double temp, double kb);
void orc_correction(int n, double* xp1, const double* fp, const int* statu, const int* ityp,
                    const double* cm, double h, double* ekin);

/* Full NVE run: [initial list+force if init], then nsteps of predictor -> list rebuild when
 * (step0+s) % nb_uptab == 0 -> half-list FS force -> corrector.  Returns ORC_OK or error. */
int orc_md_run(int n, double* xp, double* xp1, double* fp, double* dis, int* statu,
               const int* ityp, const double* cm, double h, int nsteps, int nb_uptab, int step0,
               int init, const double lo[3], const double len[3], const int ifpd[3],
               const double* nb_rm, const double* ru, int mxkvois, const orc_ftable* t,
               int* kvois, int* indi, double* epot);

/* ---------------- RNG and Maxwell velocities ---------------- */
double orc_drand32(int* seed2);
/* Philox4x32-10 (Salmon et al., SC'11): out = the block of counter ctr[4] under key[2]. */
void orc_philox4x32_10(const unsigned* ctr, const unsigned* key, unsigned* out);
/* Thermalizing_MC_DEV (MD_DiffScheme_GPU.F90:1451-1805) fed by Philox4x32-10 uniforms: atoms in
 * the caller's order, box b holding natoms[b] consecutive atoms; xp1 (N,3) column-major in/out. */
void orc_thermalize_philox(int n, int nbox, const int* natoms, const int* ityp, const int* statu,
                           const double* cm, double tkb, unsigned long long seed,
                           unsigned long long draw, double* xp1);
/* Thermalize_MC_SimMDBox over all atoms: three passes (x, y, z) over atoms in index order,
 * two DRAND32 draws per component; subtracts the mass-weighted mean velocity. kb in
 * energy/K; cm per group. */
void orc_thermalize_mc(int n, const int* ityp, const double* cm, int ngroup, double temp,
                       double kb, int* seed2, double* xp1);

/* Pairwise (cascade) sum of n doubles, the summation used for totals. */
double orc_sum(const double* a, int n);

#ifdef __cplusplus
}
#endif
#endif
