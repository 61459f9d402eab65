/*
 * msm_hip.h -- C-ABI of the MI355X-native tabulated FS/EAM molecular-dynamics force path
 * (libmsm_hip.so: hand-written gfx950 HIP kernels, FP64).
 *
 * The entry points replace the device-side call surface of the reference MDLIB:
 *   force class            MDLIB/sor/CommonGPU/MD_ForceClass_Register_GPU.F90:248-259, 363-442
 *   FS force               MDLIB/sor/CommonGPU/MD_FS_ForceTable_GPU.F90:178-257, 934-994, 1451-1700
 *   EAM force              MDLIB/sor/CommonGPU/MD_EAM_ForceTable_GPU.F90:192-278, 891-1010, 1470-1700
 *   neighbour list         MDLIB/sor/CommonGPU/MD_NeighborsList_GPU.F90:121-181, 604-693, 1347-1732
 *   integrator             MDLIB/sor/CommonGPU/MD_DiffScheme_GPU.F90:604-682, 821-847
 *   step loop              MDLIB/sor/Appshell/MD_Method_GenericMD_GPU.F90:496-659
 * and the host-side table / box-setup helpers that feed them:
 *   table generation       MDLIB/sor/Common/MD_TypeDef_ForceTable.F90:530-622, 890-1056, 1151-1230
 *   potentials             MDLIB/sor/Potentials/EM_TB_WangJun_W-HE_2010/FS_Ackland_WW.F90:23-91,
 *                          MDLIB/sor/Potentials/EM_TB_Cleri_Rosato_PRB48_22/TB_CR_AuAu.F90
 *   RNG / Maxwell          LIB/sor/f/RandGenerators/DRAND32.F90:60-85,
 *                          MDLIB/sor/Common/MD_TypeDef_SimBox.F90:4865-4942
 *
 * Array conventions are those of the reference's Fortran arrays, so Fortran callers pass their
 * arrays unchanged through ISO_C_BINDING:
 *   - XP(N,3), XP1(N,3), FP(N,3), DIS(N,3): column-major, x at [i], y at [i+N], z at [i+2N];
 *   - ITYP(N): 1-based atom group; STATU(N): CP_STATU_* bit set;
 *   - force tables POTR(NKIND,NTAB) etc. column-major: T(k,i) at [(k-1) + (i-1)*NKIND];
 *     FEMBD(NKIND1,NEMBD) likewise; KPAIR(NGROUP,NGROUP) at [(i-1) + (j-1)*NGROUP] holding a
 *     1-based kind; KEMBD(NGROUP) 1-based;
 *   - RU(NGROUP,NGROUP), NB_RM(NGROUP,NGROUP) column-major;
 *   - neighbour lists KVOIS(N), INDI(LD, MXKVOIS) slot-major with 1-based atom ids, in the
 *     caller's atom order (as Copyout_NeighboreList_B0, MD_NeighborsList_GPU.F90:604-693).
 * The library is unit-agnostic: positions, table abscissae, masses and time step must be in one
 * consistent unit system (the reference uses CGS; the tests use eV, Angstrom, fs with masses in
 * eV fs^2 / Angstrom^2).  Box shapes are orthorhombic (BOXSHAPE = identity).
 *
 * Every entry point returns MSM_OK (0) or an MSM_ERR_* code; nothing stops the process or reads
 * stdin.  msm_last_error() gives a message for the last failing call on a context.
 */
#ifndef MSM_HIP_H
#define MSM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MSM_FS_TYPE  0   /* F = -sqrt(rho) built in ("FS_TYPE") */
#define MSM_EAM_TYPE 1   /* F(rho) tabulated ("EAM_TYPE") */

/* status bits, MDLIB/sor/Common/MD_Const.F90:79-99 */
#define MSM_STATU_ACTIVE    1
#define MSM_STATU_FIXPOSX   2
#define MSM_STATU_FIXPOSY   4
#define MSM_STATU_FIXPOSZ   8
#define MSM_STATU_FIXVELX   16
#define MSM_STATU_FIXVELY   32
#define MSM_STATU_FIXVELZ   64
#define MSM_STATU_OUTOFBOX  65536
#define MSM_STATU_REFLECT   131072
#define MSM_STATU_TRANSMIT  262144
#define MSM_STATU_PASSBOUND 524288

#define MSM_OK                  0
#define MSM_ERR_ARG             1  /* invalid argument or size */
#define MSM_ERR_HIP             2  /* HIP runtime error */
#define MSM_ERR_NOMEM           3  /* device allocation failed */
#define MSM_ERR_NLIST_OVERFLOW  4  /* an atom has more than MXKVOIS neighbours */
#define MSM_ERR_OUTOFBOX        5  /* an active atom lies outside its box's cell grid */
#define MSM_ERR_EMBD_RANGE      6  /* a density fell beyond the embedding table */
#define MSM_ERR_STATE           7  /* call out of order (no table / box / atoms / list yet) */

typedef struct msm_sim msm_sim;

/* Pair/embedding callback: (r, out value, out derivative term) by reference, the shape of the
 * reference's NNFORCE/NEFORCE/EMBDF interfaces (MD_TypeDef_ForceTable.F90:46-78), so a Fortran
 * bind(C) subroutine can be passed with c_funloc. */
typedef void (*msm_pairfn)(const double* r, double* v, double* dv);

const char* msm_strerror(int code);

/* ------------------------------------------------------------------------------------------ */
/* Host-side table generation (no GPU needed)                                                  */
/* ------------------------------------------------------------------------------------------ */

/* One kind of Create_Pairwise_ForceTable (MD_TypeDef_ForceTable.F90:949-976): entries
 * i=1..ntab at r_i=(i/CSI)^2, CSI=ntab/sqrt(rmax); POTR=nn.v*r, FPOTR=nn.dv*r, POTB=ne.v,
 * FPOTB=ne.dv (ne may be NULL: zeros).  Output arrays hold ntab values with element stride
 * `stride` (use NKIND to write straight into a POTR(NKIND,NTAB) array at offset k-1).
 * rhomx_inout accumulates max(POTB*rhomult) as FTable%RHOMX does (start it at -1). */
int msm_table_pairwise(int ntab, double rmax, msm_pairfn nn, msm_pairfn ne, double rhomult,
                       double* rhomx_inout, double* potr, double* fpotr, double* potb,
                       double* fpotb, int stride);
/* One kind of Create_EMBDFUNTable (:1043-1048): F(i),F'(i) at rho=(i-1)*rhod, i=1..nembd. */
int msm_table_embd(int nembd, double rhod, msm_pairfn embd, double* fembd, double* dfembd,
                   int stride);

/* Built-in potential forms in eV / Angstrom (NN: 0.5*V and -dV/dr; NE: rho and -drho/dr;
 * EMB: F and dF/drho). */
void msm_pot_wfs_nn(const double* r, double* potr, double* fpotr);       /* benchmark W FS pair */
void msm_pot_wfs_ne(const double* r, double* potb, double* fpotb);       /* A^2 (r-4.4002)^2    */
void msm_pot_ackland_ww_nn(const double* r, double* potr, double* fpotr);/* FS_Ackland_WW.F90   */
void msm_pot_ackland_ww_ne(const double* r, double* potb, double* fpotb);
void msm_emb_fs(const double* rho, double* f, double* df);               /* -sqrt(rho)          */
/* Cleri-Rosato TB-SMA (PRB 48 (1993) 22) in the reference's form (TB_CR_AuAu.F90:17-53):
 * NN = A exp(-p(r/r0-1)), 2 NN p/r0; NE = xi^2 exp(-2q(r/r0-1)), 2 q/r0 NE; truncated only by
 * the force cutoff, as in the reference.  pair = 0:Cu-Cu 1:Ni-Ni 2:Cu-Ni (Cu, Ni from the
 * paper's table; Cu-Ni by arithmetic means of A, p, q, r0 and the geometric mean of xi). */
void msm_pot_tbsma_nn(int pair, const double* r, double* potr, double* fpotr);
void msm_pot_tbsma_ne(int pair, const double* r, double* potb, double* fpotb);
void msm_pot_tbsma_cucu_nn(const double* r, double* potr, double* fpotr);
void msm_pot_tbsma_cucu_ne(const double* r, double* potb, double* fpotb);
void msm_pot_tbsma_nini_nn(const double* r, double* potr, double* fpotr);
void msm_pot_tbsma_nini_ne(const double* r, double* potb, double* fpotb);
void msm_pot_tbsma_cuni_nn(const double* r, double* potr, double* fpotr);
void msm_pot_tbsma_cuni_ne(const double* r, double* potb, double* fpotb);

/* DRAND32 (DRAND32.F90:60-85): returns the next number and advances seed[2]. */
double msm_drand32(int* seed);
/* Thermalize_MC_SimMDBox (MD_TypeDef_SimBox.F90:4865-4942) over atoms [0,n) in index order:
 * Maxwell components v0*sqrt(-ln z1)*cos(2 pi z2), v0=sqrt(2 T kb / m), minus the mass-weighted
 * mean.  cm is per group (1-based ITYP); writes XP1(N,3). */
int msm_thermalize_mc(int n, const int* ityp, const double* cm, double temp, double kb,
                      int* seed, double* xp1);

/* ------------------------------------------------------------------------------------------ */
/* Device context                                                                              */
/* ------------------------------------------------------------------------------------------ */

/* Create a context on HIP device `device` (Initialize_DEVICES / Initialize_Globle_Variables_DEV). */
int msm_create(msm_sim** out, int device);
int msm_destroy(msm_sim* sim);
const char* msm_last_error(const msm_sim* sim);

/* Copy a force table to the device (INITIALIZE_FS_Force_Table_DEV / INITIALIZE_EAM_Force_Table_DEV).
 * pot_type MSM_FS_TYPE ignores the embedding arguments (may be 0/NULL). */
int msm_set_force_table(msm_sim* sim, int pot_type, int ngroup, int nkind, int ntab, double csi,
                        const int* kpair, const double* potr, const double* fpotr,
                        const double* potb, const double* fpotb, int nkind1, int nembd,
                        double rhod, const int* kembd, const double* fembd,
                        const double* dfembd);

/* Cutoffs: force cutoff RU(NGROUP,NGROUP) (the device kernels use max(RU^2), as the reference
 * kernels do), neighbour cutoff NB_RM(NGROUP,NGROUP), list capacity MXKVOIS, and per-group
 * masses CM(NGROUP). */
int msm_set_ctrl(msm_sim* sim, int ngroup, const double* ru, const double* nb_rm, int mxkvois,
                 const double* cm);

/* Boxes (MULTIBOX generalised to uneven sizes): nbox boxes, box b holding natoms[b] consecutive
 * atoms of the arrays given to msm_set_atoms; lo[3*b..], len[3*b..] in length units; ifpd[3*b..]
 * periodic flags. */
int msm_set_boxes(msm_sim* sim, int nbox, const int* natoms, const double* lo, const double* len,
                  const int* ifpd);

/* Upload atoms (copy-in of SimBox XP/XP1/ITYP/STATU; DIS and FP reset to zero). */
int msm_set_atoms(msm_sim* sim, int n, const double* xp, const double* xp1, const int* ityp,
                  const int* statu);

/* Copy-out in the caller's atom order; any pointer may be NULL. */
int msm_get_atoms(msm_sim* sim, double* xp, double* xp1, double* fp, int* statu, double* dis);
/* CopyXPFrom_Host_to_Devices / CopyXP1From_Host_to_Devices (MD_Globle_Variables_GPU.F90:349-373):
 * positions and/or velocities (either may be NULL) of all atoms, (N,3) column-major in the
 * caller's order, written into the device state without a list rebuild (the caller keeps the
 * moves within what the list allows, as the reference's quench loops do). */
int msm_put_atoms(msm_sim* sim, const double* xp, const double* xp1);
/* External force and energy terms (AddExtForce_ForceClass / Cumulate_ExtForce,
 * MD_ForceClass_Register_GPU.F90:584-680, 982-1180): the reference calls the registered
 * ForceArray/EpotArray routines after the force (energy) pass and they add their terms to the
 * device arrays.  msm_add_forces adds FEXT(N,3) (column-major, caller's atom order) to the current
 * forces, msm_add_epot adds EEXT(N) to the per-atom energies (and renews the box totals of
 * msm_get_box_epot); on_device != 0 means the array is a device pointer of this context's device.
 * The fused steps of msm_run compute their forces inside the graph and do not see these terms:
 * step with msm_predictor, msm_calc_force, msm_add_forces, msm_correction instead. */
int msm_add_forces(msm_sim* sim, const double* fext, int on_device);
int msm_add_epot(msm_sim* sim, const double* eext, int on_device);
/* DIS(N,3) in the caller's order written into the device state (the restart of a run:
 * msm_set_atoms resets DIS to zero). */
int msm_put_dis(msm_sim* sim, const double* dis);
/* F'(rho) per atom (the reference's DEN array) and per-atom potential energy EPOT. */
int msm_get_den(msm_sim* sim, double* den);
int msm_get_epot(msm_sim* sim, double* epot);
/* Total potential energy per box (pairwise tree sum on the device); valid after msm_calc_epot. */
int msm_get_box_epot(msm_sim* sim, double* ebox);

/* Cal_NeighBoreList_DEV: device cell binning + deterministic sort + 27-cell full list, FP64
 * distances with the reference host predicate (strict <, shifted images xj+L). */
int msm_build_neighbors(msm_sim* sim);
/* Copyout_NeighboreList: KVOIS(N) and INDI(ld, MXKVOIS) (1-based ids) in caller order. */
int msm_get_neighbors(msm_sim* sim, int* kvois, int* indi, int ld);

/* CALFORCE_*_Force_Table2A_DEV: density pass then force pass. */
int msm_calc_force(msm_sim* sim);
/* CALDEN_FS/EAM_Force_Table2A_DEV (MD_FS_ForceTable_GPU.F90:875-930, bound as pCalEDen at
 * MD_ForceClass_Register_GPU.F90:405,416): the density pass alone -- F'(rho) of every atom is left
 * on the device (msm_get_den copies it out); forces are not touched. */
int msm_calc_density(msm_sim* sim);
/* CALEPOT_*_Force_Table2A_DEV: per-atom potential energy (and box totals). */
int msm_calc_epot(msm_sim* sim);
/* Predictor_DEV / Correction_DEV (Predictor_KERNEL0 / Correction_KERNEL semantics). */
/* Virial tensor per box (CALPTENSOR_FS/EAM_Force_Table2A_DEV, MD_FS_ForceTable_GPU.F90:998-1447):
 * sum over pairs of s (x) s * FORTOT (each pair once), VTENSOR(3,3) column-major at vtensor[9*b],
 * b = 0..nbox-1; recomputes the density pass and stores the forces as msm_calc_force does.  The
 * reference's COPYOUT_VIRIALTENSOR returns the mean over boxes; callers average these. */
int msm_calc_ptensor(msm_sim* sim, double* vtensor);
/* msm_calc_ptensor without any device allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (serving loops, harnesses that forbid hipMalloc).  Same
 * inputs, outputs, forces and flags as msm_calc_ptensor (the per-box sums are added in another
 * fixed order, so the tensor agrees to rounding, the forces bit for bit).  Its partial sums live
 * in a fixed 4 MiB device array per GPU, held from the first msm_create on that GPU until the
 * process ends: MSM_ERR_ARG when nbox * ceil(largest box's atoms / 256) * 6 + 9 * nbox exceeds
 * 524288 doubles (one box of up to about 22 million atoms).  Calls on one GPU share the array:
 * calls that overlap in time on the same GPU (other contexts, other threads) race. */
int msm_calc_ptensor_noalloc(msm_sim* sim, double* vtensor);

/* Per-atom stress (virial part), the reference's CAL_FS/EAM_AtomicStress_KERNEL through
 * Cal_AtomicStress (MD_FS_ForceTable_GPU.F90:1746-2076, Analysis/AtomicVirialStress_GPU.F90:734-773):
 * for every active atom i, P_ab = sum over its full list of s_a s_b FORTOT (no factor 1/2; zero
 * for inactive atoms), written as AVP(N,9) column-major in the caller's atom order with components
 * P11, P12, P13, P21, P22, P23, P31, P32, P33 at avp[c*N + i].  Recomputes the density pass and
 * stores the forces as msm_calc_force does; the sum over atoms is twice msm_calc_ptensor's. */
int msm_calc_atomic_stress(msm_sim* sim, double* avp);

/* CheckTimestep_DEV (MD_DiffScheme_GPU.F90:1068-1210): *exceeded = 1 when an active atom's
 * displacement over a step th, |th v + th^2/2 F/m| (components without a FIXPOS bit), exceeds
 * sqrt(mxd2); else 0.  The variable-step search of Predictor_DEV (:629-657) halves th from HMX
 * until this flag is 0. */
int msm_check_timestep(msm_sim* sim, double th, double mxd2, int* exceeded);

/* Thermalisation and quench around the force path (MD_DiffScheme_GPU.F90).
 * msm_damping: DAMPING_KERNEL (:125-187) -- each velocity component opposite to its force
 * component is set to zero, and components with FIXPOS set (every atom).
 * msm_calc_ekin: CALEKIN_KERNEL (:851-904) and Cal_GlobalT_DEV (:1042-1055) per box -- ekin (may be
 * NULL) receives CM/2 |v|^2 in the caller's order for active atoms with no FIXPOS bit and -1e32
 * for the others; tbox (may be NULL) the temperature 2 <EK> / (3 kb) of each box over the atoms
 * that count (kb in the caller's energy unit per K).
 * msm_scale_velocities: VelScaling_DEV (:1262-1448) -- velocities of active atoms times
 * sqrt(temp * 3 kb / 2 / <EK>_box), components with FIXPOS or FIXVEL set to zero; MSM_ERR_ARG if a
 * box has no kinetic energy.
 * msm_dyn_damp: Do_DynDamp_Forsteps_DEV (:1809-1883) -- up to maxsteps of damping, predictor,
 * force, per-atom energy, stopping when max |EPOT_i - EPOT_i(previous)| <= mindele (before the
 * corrector), else corrector; the list is not rebuilt.  steps receives the iteration it stopped
 * at (maxsteps + 1 if it did not converge), delepot the last max |dEPOT|.
 * msm_calc_ekin, msm_scale_velocities and msm_dyn_damp need a built neighbour list (the atoms sit
 * in cell order on the device from the first msm_build_neighbors on). */
int msm_damping(msm_sim* sim);
/* Thermalizing_MC_DEV (MD_DiffScheme_GPU.F90:1451-1805): every active atom's velocity components
 * whose FIXVEL and FIXPOS bits are clear become V0 sqrt(-ln z1) cos(2 pi z2), V0 = sqrt(2 temp kb
 * / CM) (the other components keep their value), inactive atoms get zero velocity, then each
 * box's mass-weighted mean velocity over all its atoms is subtracted from all its atoms.  The
 * uniforms z1, z2 in (0, 1] come from Philox4x32-10 (Salmon et al., SC'11) keyed by `seed`, with
 * counter (caller's atom index, component, draw), draw = the number of earlier msm_thermalize calls
 * on the context -- the reference's cuRAND XORWOW stream is not reproduced, the distribution is.
 * Needs a built neighbour list (cell order). */
int msm_thermalize(msm_sim* sim, double temp, double kb, unsigned long long seed);
int msm_calc_ekin(msm_sim* sim, double kb, double* ekin, double* tbox);
int msm_scale_velocities(msm_sim* sim, double temp, double kb);
int msm_dyn_damp(msm_sim* sim, int maxsteps, double mindele, double h, int* steps,
                 double* delepot);
/* Steepest-descent quench on the current neighbour list (MD_SteepestScheme_GPU.F90), lengths in
 * the caller's unit (the reference passes STEEPEST_MxStep and STEEPEST_MiStep times the lattice
 * unit RR), energies in the caller's unit.
 * linesearch = 0: Do_Steepest0_Forsteps_DEV (:20-152) -- moves of alpha_k F, alpha_k from
 *   <dx, dF> / <dF, dF> of the last two steps (alpha when negative), the largest component capped
 *   at maxdis; stops when that largest component (before the cap) is <= mindis or the largest
 *   per-atom energy change is <= mindele.  steps receives the iteration it stopped at (0 when the
 *   first move is already below mindis, maxsteps + 1 if it did not converge).
 * linesearch = 1: Do_Steepest1_Forsteps_DEV (:156-260) -- along the normalised force, secant
 *   steps on F . D (capped at maxdis) until a step is <= mindis, repeated until the per-atom
 *   energy change of a line search is <= mindele (the reference uses 1e-3 eV); steps receives the
 *   number of force evaluations of the line searches (> maxsteps if it did not converge).
 * Positions move by the displacement with the bounds of DevAdd (periodic directions wrap once);
 * atoms are moved whatever their FIXPOS bits, as in the reference.  delepot receives the last
 * max |dEPOT|, maxmove the last largest displacement component (or line-search step). */
int msm_steepest(msm_sim* sim, int maxsteps, int linesearch, double alpha, double maxdis,
                 double mindis, double mindele, int* steps, double* delepot, double* maxmove);
/* Conjugate-gradient quench on the current neighbour list (MD_CGScheme_GPU.F90), directions
 * updated by Polak-Ribiere, d = F + (F . (F - F0) / |F0|^2) d, from d = F; lengths and energies as
 * for msm_steepest; stops when |F0|^2 <= 1e-64 (steps = 0 at the start).
 * linesearch = 0: Do_CG0_Forsteps_DEV (:16-146) -- per direction a trial step of maxdis/|F0| d
 *   and one secant step; stops when the per-atom energy change is <= mindele; steps receives the
 *   iteration (maxsteps + 1 if it did not converge).
 * linesearch = 1: Do_CG1_Forsteps_DEV (:150-291) -- secant steps along each direction until a
 *   step is <= mindis; steps counts the moves (> maxsteps if it did not converge). */
int msm_cg(msm_sim* sim, int maxsteps, int linesearch, double maxdis, double mindis,
           double mindele, int* steps, double* delepot);
int msm_predictor(msm_sim* sim, double h);
int msm_correction(msm_sim* sim, double h);

/* nsteps of For_One_Step on the device: predictor, list rebuild when (step0+s) % nb_uptab == 0,
 * density, force, corrector.  The corrector of step s and the predictor of step s+1 run fused in
 * the force kernel; the sequence between rebuilds is replayed from a captured HIP graph.
 * Requires forces of the current positions (msm_calc_force) on entry. */
int msm_run(msm_sim* sim, int nsteps, int nb_uptab, int step0, double h);

/* Wait for all device work of the context. */
int msm_sync(msm_sim* sim);

/* Neighbour-list statistics of the last build: total entries (sum of KVOIS) and max KVOIS. */
int msm_list_stats(msm_sim* sim, long long* total, int* maxk);

/* Options.  MSM_OPT_TRACK_DIS (default 1): accumulate DIS in the fused predictor of msm_run.
 * MSM_OPT_KERNEL_TIMING (default 0): msm_run enqueues its kernels directly (no graph replay) with
 * HIP events around each launch and accumulates per-class device time. */
#define MSM_OPT_TRACK_DIS     1
#define MSM_OPT_KERNEL_TIMING 2
#define MSM_OPT_NLIST_KERNEL  3  /* 0 (default): LDS-staged cell bricks, small or wide by cell
                                    occupancy; 1: per-atom global scan; 2: small bricks (one lane
                                    per atom); 3: wide bricks (several lanes per atom) */
#define MSM_OPT_FORCE_PATH    4  /* density/force passes: 0 LDS bricks; 1, 2, 4, 8: global list
                                    with that many lanes per atom (default 2); 3: two lanes per
                                    atom, each 32-byte record gathered by a lane pair; 5: staged
                                    bricks (a brick's halo staged in LDS once, pairs read as halo
                                    slots; needs a list from the brick list kernels, else as 2) */
#define MSM_OPT_COMPACT_LIST  5  /* 0 (default); 1: on the global-list path with 1, 2 or 4 lanes
                                    per atom,
                                    the density pass records the pairs inside the force cutoff and
                                    the force pass visits only those */
#define MSM_OPT_TABLE_LDS     6  /* 1 (default): the passes read the table's knots from a
                                    window held in LDS (persistent grid) -- on the global-list
                                    paths 1, 2, 3, 4 (no compaction) and, when the list was
                                    built by the LDS-brick kernels, on path 0 (then positions
                                    and knots both come from LDS) */
/* Shape of the LDS-table passes: threads per workgroup (256, 512, 1024; default 1024) and
 * workgroups per CU (1..16, default 1; capped by what fits); the LDS window is the CU's LDS
 * divided by the workgroups per CU. */
#define MSM_OPT_LT_DENSITY_BLOCK 7
#define MSM_OPT_LT_DENSITY_WGS   8
#define MSM_OPT_LT_FORCE_BLOCK   9
#define MSM_OPT_LT_FORCE_WGS     10
/* Cell numbering (and so the order of the cell-sorted atoms): bricks of value^3 cells are
 * contiguous (default 1 = row-major).  Takes effect at the next msm_build_neighbors, which is
 * then required. */
#define MSM_OPT_CELL_BRICK       11
/* Lanes per home atom (1 default, 2 or 4) of the LDS-brick passes with the table in LDS. */
#define MSM_OPT_BRICK_LANES      12
int msm_set_option(msm_sim* sim, int option, int value);
/* Lanes per atom (1, 2, 4) of the LDS-table force pass when it should differ from the density
 * pass's (MSM_OPT_FORCE_PATH); 0 (default) follows MSM_OPT_FORCE_PATH. */
#define MSM_OPT_FORCE_LANES      14
/* Chunk order of the persistent LDS-table passes: 0 (default) each workgroup walks a contiguous
 * range of its XCD's atom chunks; 1 the workgroups of an XCD take its chunks round-robin, so the
 * chunks in flight on one XCD are adjacent and share that XCD's L2. */
#define MSM_OPT_CHUNK_SWEEP      15
/* Read-only: the density/force kernels the current options and list select (MSM_PASS_*). */
#define MSM_OPT_PASS_KERNEL      13
#define MSM_PASS_BRICKS      0  /* LDS-staged positions, global table */
#define MSM_PASS_GLOBAL      1  /* one lane per atom, global list and table */
#define MSM_PASS_LANES       2  /* several lanes per atom, global list and table */
#define MSM_PASS_LANE_PAIR   3  /* two lanes per atom, lane-pair gathers */
#define MSM_PASS_LDS_TABLE   4  /* global list, table window in LDS */
#define MSM_PASS_BRICK_TABLE 5  /* LDS-staged positions and table window in LDS */
#define MSM_PASS_STAGED      6  /* staged bricks: halo images resolved once, slots per pair */
/* Current value of an option (every MSM_OPT_* above). */
int msm_get_option(msm_sim* sim, int option, int* value);

/* Adjoint (reverse mode) of the force and energy evaluation with respect to the table knots,
 * for every table kind.  Given dL/dF (dldf, 3*N doubles in the caller's atom order and FP(N,3)
 * layout) and dL/dE per box (dlde, one double per box, E_b = the box sum of msm_get_epot), writes
 * dL/d(knot) for every kind k and knot i = 1..NTAB of POTR, FPOTR, POTB, FPOTB, at
 * [k + (i-1)*NKIND] as msm_set_force_table reads them, and, for EAM tables, for every embedding
 * kind and knot i = 1..NEMBD of FEMBD and DFEMBD at [k + (i-1)*NKIND1] (may be NULL for FS
 * tables).  With several atom types a pair (i, j) reaches the kinds KPAIR(ti, tj) (density,
 * pair force, pair energy, and the FPOTB term of F'_i) and KPAIR(tj, ti) (the FPOTB term of
 * F'_j); kinds no pair reaches get zeros.  Uses the current positions and neighbour list;
 * recomputes rho and F' itself.  The per-pair terms are summed with LDS atomics, so repeated
 * calls agree to rounding.  Corresponds to the C-ABI item "msm_adjoint(h, dLdF, dLdE,
 * out_grad_tables)" of the hot-path boundary; the reference has no derivative code. */
int msm_adjoint(msm_sim* sim, const double* dldf, const double* dlde, double* g_potr,
                double* g_fpotr, double* g_potb, double* g_fpotb, double* g_fembd,
                double* g_dfembd);

/* Device-resident force matching (no counterpart in the reference, which has no derivative
 * code; for fitting loops that keep every per-atom array on the device).  The potential is a
 * cubic B-spline model of one kind: coef holds 3 rows of ld doubles (row-major), the pair row
 * u(s), the density row rho(s) (s = sqrt(r)*CSI, knots 1..NTAB) and the embedding row F(m)
 * (m = rho/RHOD + 1, knots 1..NEMBD; ignored for FS tables); coefficients beyond NTAB/NEMBD are
 * not read.  At knot i a row c has value v = ((c[i-2] + 4 c[i-1]) + c[i]) / 6 and slope
 * s = (c[i] - c[i-2]) / 2 (c zero outside the row), and the tables are
 *   POTR = v(u) (i/CSI)^2, FPOTR = -i s(u), POTB = v(rho), FPOTB = ((-s(rho)) CSI^2) / (2i),
 *   FEMBD = v(F), DFEMBD = s(F) / RHOD.
 *
 * msm_set_spline_table builds those tables on the device into the buffers of the table set
 * by msm_set_force_table (whose NTAB, CSI, NEMBD, RHOD, KPAIR and KEMBD it keeps; it must hold
 * one pair kind and, for EAM, one embedding kind; ld >= max(NTAB, NEMBD)).  The result is
 * bit-identical to msm_set_force_table on the same tables evaluated on the host with these
 * formulas.  coef is a device pointer when coef_on_device is nonzero (read before the call
 * returns); host coefficients are staged, and the build is queued on the context's stream
 * ahead of every later call (msm_set_force_table waits for it).
 *
 * msm_fm_set_targets keeps the target forces (3*N doubles, caller order, FP(N,3) layout) and box
 * energies (one per box) on the device with the energy weight w_e >= 0.
 *
 * msm_fm_loss_grad evaluates the forces and box energies with the current tables and returns
 *   L = sum_i |F_i - F*_i|^2 / N + w_e sum_b ((E_b - E*_b) / N_b)^2 / NBOX
 * in *loss and, when coef_grad is not NULL, dL/dcoef (3 rows of ld, zero beyond NTAB/NEMBD)
 * through the adjoint passes of msm_adjoint and the transpose of the table map above
 * (coef_grad is a device pointer when grad_on_device is nonzero).  Only the loss (and the
 * gradient, when on the host) cross PCIe.  Sums run in a fixed order except the adjoint's LDS
 * atomics, so the gradient of repeated calls agrees to rounding.  Repeated calls replay one
 * captured HIP graph of the whole evaluation (recaptured when the atoms, list, tables' layout,
 * options, ld, w_e or the gradient pointer change; with MSM_OPT_KERNEL_TIMING the kernels are
 * launched directly and timed). */
int msm_set_spline_table(msm_sim* sim, const double* coef, int ld, int coef_on_device);
int msm_fm_set_targets(msm_sim* sim, const double* f_target, const double* e_target, double w_e);
int msm_fm_loss_grad(msm_sim* sim, double* loss, double* coef_grad, int ld, int grad_on_device);

/* Kernel-variant autotuner (no counterpart in the reference, whose kernel shapes are fixed at
 * compile time).  On the current atoms and list it times `reps` rebuilds with each list-build
 * kernel and `reps` density + force passes (no integration) of every applicable variant of the
 * density/force kernels (lanes per atom, LDS table and its workgroup shape, lane-pair gather,
 * LDS bricks), sets the options to the fastest, and recomputes F' and F with it.  choice[8]
 * (may be NULL) receives {force path, table LDS, density block, density workgroups per CU, force
 * block, force workgroups per CU, list kernel (MSM_OPT_NLIST_KERNEL value), brick lanes
 * (MSM_OPT_BRICK_LANES value)}; best_ms[2] (may
 * be NULL) the chosen density+force and rebuild times in ms.  Positions and velocities are
 * unchanged. */
int msm_autotune(msm_sim* sim, int reps, int* choice, double* best_ms);

/* Kernel classes for msm_kernel_times. */
#define MSM_KCLASS_DENSITY   0
#define MSM_KCLASS_FORCE     1  /* force pass with the fused corrector/predictor */
#define MSM_KCLASS_REBUILD   2  /* whole neighbour-list rebuild (binning, sort, gather, list) */
#define MSM_KCLASS_PREDICTOR 3
#define MSM_KCLASS_NLIST     4  /* the list kernel alone (inside the rebuild) */
#define MSM_KCLASS_ADJOINT   5  /* the device passes of msm_adjoint (rho, pair, density, sums) */
#define MSM_KCLASS_COUNT     6
/* Accumulated device milliseconds and launch counts per kernel class since the last reset
 * (collected while MSM_OPT_KERNEL_TIMING is on); reset != 0 clears the counters after reading. */
int msm_kernel_times(msm_sim* sim, double* ms, long long* count, int nclass, int reset);

#ifdef __cplusplus
}
#endif
#endif
