!  MD_ForceLib_HIP.F90 -- Fortran drop-in for the MDLIB call surface of the MD force hot path:
!  force tables (MD_TypeDef_ForceTable.F90), simulation box and control blocks (subset of
!  MD_TypeDef_SimBox.F90 / MD_TypeDef_SimCtrlParam.F90), the force class (MD_ForceClass_Register_GPU.F90)
!  and the device routines of the step (MD_Globle_Variables_GPU.F90, MD_NeighborsList_GPU.F90,
!  MD_DiffScheme_GPU.F90, MD_Method_GenericMD_GPU.F90).  Every device operation goes through the
!  C-ABI of libmsm_hip.so (include/msm_hip.h) by ISO_C_BINDING; no Fortran code runs on the GPU.
!
!  Errors follow the reference's convention: "MDPSCU Error: ..." on the output unit and STOP,
!  unless the caller passes the optional IERR argument, which then receives the C-ABI code.
!
!  Units are the caller's (the library is unit-agnostic): lengths, energies, masses and the time
!  step must be consistent (e.g. eV, Angstrom, eV fs^2/A^2, fs).

module MSM_HIP_C
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: MSM_OK = 0, MSM_ERR_ARG = 1
  integer(c_int), parameter :: MSM_FS_TYPE = 0, MSM_EAM_TYPE = 1
  integer(c_int), parameter :: MSM_OPT_TRACK_DIS = 1, MSM_OPT_KERNEL_TIMING = 2, &
                               MSM_OPT_NLIST_KERNEL = 3, MSM_OPT_FORCE_PATH = 4

  interface
    function msm_strerror(code) bind(C, name="msm_strerror") result(p)
      import :: c_int, c_ptr
      integer(c_int), value :: code
      type(c_ptr) :: p
    end function msm_strerror

    function msm_create(sim, device) bind(C, name="msm_create") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: sim
      integer(c_int), value :: device
      integer(c_int) :: rc
    end function msm_create

    function msm_destroy(sim) bind(C, name="msm_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int) :: rc
    end function msm_destroy

    function msm_last_error(sim) bind(C, name="msm_last_error") result(p)
      import :: c_ptr
      type(c_ptr), value :: sim
      type(c_ptr) :: p
    end function msm_last_error

    function msm_set_force_table(sim, pot_type, ngroup, nkind, ntab, csi, kpair, potr, fpotr, &
                                 potb, fpotb, nkind1, nembd, rhod, kembd, fembd, dfembd) &
        bind(C, name="msm_set_force_table") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: pot_type, ngroup, nkind, ntab, nkind1, nembd
      real(c_double), value :: csi, rhod
      integer(c_int), intent(in) :: kpair(*)
      real(c_double), intent(in) :: potr(*), fpotr(*), potb(*), fpotb(*)
      type(c_ptr), value :: kembd, fembd, dfembd
      integer(c_int) :: rc
    end function msm_set_force_table

    function msm_set_ctrl(sim, ngroup, ru, nb_rm, mxkvois, cm) bind(C, name="msm_set_ctrl") &
        result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: ngroup, mxkvois
      real(c_double), intent(in) :: ru(*), nb_rm(*), cm(*)
      integer(c_int) :: rc
    end function msm_set_ctrl

    function msm_set_boxes(sim, nbox, natoms, lo, len, ifpd) bind(C, name="msm_set_boxes") &
        result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: nbox
      integer(c_int), intent(in) :: natoms(*), ifpd(*)
      real(c_double), intent(in) :: lo(*), len(*)
      integer(c_int) :: rc
    end function msm_set_boxes

    function msm_set_atoms(sim, n, xp, xp1, ityp, statu) bind(C, name="msm_set_atoms") &
        result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: n
      real(c_double), intent(in) :: xp(*), xp1(*)
      integer(c_int), intent(in) :: ityp(*), statu(*)
      integer(c_int) :: rc
    end function msm_set_atoms

    function msm_get_atoms(sim, xp, xp1, fp, statu, dis) bind(C, name="msm_get_atoms") &
        result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim, xp, xp1, fp, statu, dis
      integer(c_int) :: rc
    end function msm_get_atoms

    function msm_check_timestep(sim, th, mxd2, exceeded) bind(C, name="msm_check_timestep") &
        result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: th, mxd2
      integer(c_int), intent(out) :: exceeded
      integer(c_int) :: rc
    end function msm_check_timestep

    function msm_put_atoms(sim, xp, xp1) bind(C, name="msm_put_atoms") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim, xp, xp1
      integer(c_int) :: rc
    end function msm_put_atoms

    function msm_get_epot(sim, epot) bind(C, name="msm_get_epot") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), intent(out) :: epot(*)
      integer(c_int) :: rc
    end function msm_get_epot

    function msm_get_den(sim, den) bind(C, name="msm_get_den") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), intent(out) :: den(*)
      integer(c_int) :: rc
    end function msm_get_den

    function msm_get_box_epot(sim, ebox) bind(C, name="msm_get_box_epot") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), intent(out) :: ebox(*)
      integer(c_int) :: rc
    end function msm_get_box_epot

    function msm_build_neighbors(sim) bind(C, name="msm_build_neighbors") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int) :: rc
    end function msm_build_neighbors

    function msm_get_neighbors(sim, kvois, indi, ld) bind(C, name="msm_get_neighbors") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), intent(out) :: kvois(*), indi(*)
      integer(c_int), value :: ld
      integer(c_int) :: rc
    end function msm_get_neighbors

    function msm_calc_force(sim) bind(C, name="msm_calc_force") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int) :: rc
    end function msm_calc_force

    function msm_calc_density(sim) bind(C, name="msm_calc_density") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int) :: rc
    end function msm_calc_density

    function msm_calc_epot(sim) bind(C, name="msm_calc_epot") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int) :: rc
    end function msm_calc_epot

    function msm_calc_ptensor(sim, vtensor) bind(C, name="msm_calc_ptensor") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), intent(out) :: vtensor(*)
      integer(c_int) :: rc
    end function msm_calc_ptensor

    function msm_calc_atomic_stress(sim, avp) bind(C, name="msm_calc_atomic_stress") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), intent(out) :: avp(*)
      integer(c_int) :: rc
    end function msm_calc_atomic_stress

    function msm_thermalize(sim, temp, kb, seed) bind(C, name="msm_thermalize") result(rc)
      import :: c_int, c_double, c_ptr, c_long_long
      type(c_ptr), value :: sim
      real(c_double), value :: temp, kb
      integer(c_long_long), value :: seed
      integer(c_int) :: rc
    end function msm_thermalize

    function msm_add_forces(sim, fext, on_device) bind(C, name="msm_add_forces") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim, fext
      integer(c_int), value :: on_device
      integer(c_int) :: rc
    end function msm_add_forces

    function msm_add_epot(sim, eext, on_device) bind(C, name="msm_add_epot") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim, eext
      integer(c_int), value :: on_device
      integer(c_int) :: rc
    end function msm_add_epot

    function msm_put_dis(sim, dis) bind(C, name="msm_put_dis") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim, dis
      integer(c_int) :: rc
    end function msm_put_dis

    function msm_damping(sim) bind(C, name="msm_damping") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int) :: rc
    end function msm_damping

    function msm_calc_ekin(sim, kb, ekin, tbox) bind(C, name="msm_calc_ekin") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: kb
      real(c_double), intent(out) :: ekin(*), tbox(*)
      integer(c_int) :: rc
    end function msm_calc_ekin

    function msm_scale_velocities(sim, temp, kb) bind(C, name="msm_scale_velocities") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: temp, kb
      integer(c_int) :: rc
    end function msm_scale_velocities

    function msm_dyn_damp(sim, maxsteps, mindele, h, steps, delepot) &
        bind(C, name="msm_dyn_damp") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: maxsteps
      real(c_double), value :: mindele, h
      integer(c_int), intent(out) :: steps
      real(c_double), intent(out) :: delepot
      integer(c_int) :: rc
    end function msm_dyn_damp

    function msm_steepest(sim, maxsteps, linesearch, alpha, maxdis, mindis, mindele, steps, &
        delepot, maxmove) bind(C, name="msm_steepest") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: maxsteps, linesearch
      real(c_double), value :: alpha, maxdis, mindis, mindele
      integer(c_int), intent(out) :: steps
      real(c_double), intent(out) :: delepot, maxmove
      integer(c_int) :: rc
    end function msm_steepest

    function msm_cg(sim, maxsteps, linesearch, maxdis, mindis, mindele, steps, delepot) &
        bind(C, name="msm_cg") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: maxsteps, linesearch
      real(c_double), value :: maxdis, mindis, mindele
      integer(c_int), intent(out) :: steps
      real(c_double), intent(out) :: delepot
      integer(c_int) :: rc
    end function msm_cg

    function msm_predictor(sim, h) bind(C, name="msm_predictor") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: h
      integer(c_int) :: rc
    end function msm_predictor

    function msm_correction(sim, h) bind(C, name="msm_correction") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: h
      integer(c_int) :: rc
    end function msm_correction

    function msm_run(sim, nsteps, nb_uptab, step0, h) bind(C, name="msm_run") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: nsteps, nb_uptab, step0
      real(c_double), value :: h
      integer(c_int) :: rc
    end function msm_run

    function msm_set_option(sim, option, value) bind(C, name="msm_set_option") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: option, value
      integer(c_int) :: rc
    end function msm_set_option

    function msm_autotune(sim, reps, choice, best_ms) bind(C, name="msm_autotune") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: reps
      integer(c_int), intent(out) :: choice(*)
      real(c_double), intent(out) :: best_ms(*)
      integer(c_int) :: rc
    end function msm_autotune

    function msm_adjoint(sim, dldf, dlde, g_potr, g_fpotr, g_potb, g_fpotb, g_fembd, g_dfembd) &
        bind(C, name="msm_adjoint") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), intent(in) :: dldf(*), dlde(*)
      real(c_double), intent(out) :: g_potr(*), g_fpotr(*), g_potb(*), g_fpotb(*)
      type(c_ptr), value :: g_fembd, g_dfembd
      integer(c_int) :: rc
    end function msm_adjoint

    function msm_drand32(seed) bind(C, name="msm_drand32") result(z)
      import :: c_int, c_double
      integer(c_int), intent(inout) :: seed(2)
      real(c_double) :: z
    end function msm_drand32

    function msm_thermalize_mc(n, ityp, cm, temp, kb, seed, xp1) &
        bind(C, name="msm_thermalize_mc") result(rc)
      import :: c_int, c_double
      integer(c_int), value :: n
      integer(c_int), intent(in) :: ityp(*)
      real(c_double), intent(in) :: cm(*)
      real(c_double), value :: temp, kb
      integer(c_int), intent(inout) :: seed(2)
      real(c_double), intent(out) :: xp1(*)
      integer(c_int) :: rc
    end function msm_thermalize_mc
  end interface

contains

  ! NUL-terminated C string -> Fortran string
  function msm_cstring(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: a(:)
    integer :: n
    if (.not. c_associated(p)) then
      s = ""
      return
    end if
    call c_f_pointer(p, a, [4096])
    n = 0
    do while (n < 4096)
      if (a(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: s)
    s = transfer(a(1:n), s)
  end function msm_cstring

end module MSM_HIP_C


module MD_ForceLib_HIP
  use iso_c_binding
  use MSM_HIP_C
  implicit none
  private

  integer, parameter, public :: KINDDF = c_double
  real(KINDDF), parameter, public :: CP_KB_EV = 8.617333262d-5   ! Boltzmann constant, eV/K
  ! the reference's CGS constants (MSMLIB/sor/Common/MSM_Const.F90:82-85)
  real(KINDDF), parameter, public :: CP_KB_CGS = 1.38054d-16       ! Boltzmann constant, erg/K
  real(KINDDF), parameter, public :: CP_EVERG = 1.60219d-12        ! erg per eV
  integer, parameter, public :: mp_MXGROUP = 16                    ! MD_Const.F90:78
  integer, parameter, public :: mp_MXEXTFORCE = 16
  ! quench schemes (MD_Const.F90:23-27)
  integer, parameter, public :: CP_DAMPSCHEME_LBFGS = 0, CP_DAMPSCHEME_DYN = 1, &
                                CP_DAMPSCHEME_CG = 2, CP_DAMPSCHEME_ST = 3, &
                                CP_DAMPSCHEME_LSEARCH = 65536
  integer, parameter :: mp_MXFORCE = 64

  ! --- MD_TypeDef_ForceTable.F90:46-78: the registered potential routines
  abstract interface
    subroutine NNFORCE_PROC(R, POTR, FPOTR)
      import :: KINDDF
      real(KINDDF), intent(in) :: R
      real(KINDDF), intent(out) :: POTR, FPOTR
    end subroutine NNFORCE_PROC
    subroutine NEFORCE_PROC(R, POTB, FPOTB)
      import :: KINDDF
      real(KINDDF), intent(in) :: R
      real(KINDDF), intent(out) :: POTB, FPOTB
    end subroutine NEFORCE_PROC
    subroutine EMBDF_PROC(RHO, FRHO, DFRHO)
      import :: KINDDF
      real(KINDDF), intent(in) :: RHO
      real(KINDDF), intent(out) :: FRHO, DFRHO
    end subroutine EMBDF_PROC
  end interface
  public :: NNFORCE_PROC, NEFORCE_PROC, EMBDF_PROC

  type :: ForceProcEntry
    logical :: REG = .false.
    procedure(NNFORCE_PROC), pointer, nopass :: pNN => null()
    procedure(NEFORCE_PROC), pointer, nopass :: pNE => null()
    procedure(EMBDF_PROC), pointer, nopass :: pEMB => null()
    character(len=256) :: NOTE = ""
    character(len=8) :: SYMB = ""   ! "TGT<-SOR" of the symbol form, upper case
  end type ForceProcEntry

  ! --- MDForceTable (MD_TypeDef_ForceTable.F90:117-155); arrays (NKIND, NTAB) / (NKIND1, NEMBD)
  type, public :: MDForceTable
    character(len=16) :: PotType = "FS_TYPE"
    integer :: NKIND = 0, NTAB = 10000, NKIND1 = 0, NEMBD = 10000
    real(KINDDF) :: Rmax = -1.0d0, RmaxSQRT = 0.0d0, CSI = 0.0d0, CSIV = 0.0d0
    real(KINDDF) :: RHOMX = -1.0d0, RHOD = 1.0d0, RHOMULT = 20.0d0
    integer, allocatable :: FPAIR(:), FPAIR1(:), KPAIR(:, :), KEMBD(:)
    real(KINDDF), allocatable :: POTR(:, :), FPOTR(:, :), POTB(:, :), FPOTB(:, :)
    real(KINDDF), allocatable :: FEMBD(:, :), DFEMBD(:, :)
    type(ForceProcEntry) :: PROC(mp_MXFORCE)
  end type MDForceTable

  ! --- SimMDBox (the fields of MD_TypeDef_SimBox.F90:43-132 that the force path, the box setup
  !     and the binary restart records read; CM and PTYPE hold NGROUP entries here and are padded
  !     to mp_MXGROUP in the records)
  type, public :: SimMDBox
    integer :: NPRT = 0, NGROUP = 1
    real(KINDDF) :: ZL(3) = 0.0d0, BOXLOW(3) = 0.0d0
    real(KINDDF) :: RR = 1.0d0   ! lattice unit, in the caller's length unit
    real(KINDDF), allocatable :: XP(:, :), XP1(:, :), FP(:, :), DIS(:, :), EPOT(:), CM(:)
    real(KINDDF), allocatable :: EKIN(:)
    integer, allocatable :: ITYP(:), STATU(:), PTYPE(:, :)
    character(len=8) :: SYMB(mp_MXGROUP) = ""
    real(KINDDF) :: LATT(3) = 0.0d0
    real(KINDDF) :: BOXSHAPE(3, 3) = reshape([1.0d0, 0.0d0, 0.0d0, 0.0d0, 1.0d0, 0.0d0, &
                                              0.0d0, 0.0d0, 1.0d0], [3, 3])
    character(len=12) :: POTTYPE = ""
    character(len=128) :: POTLIBNAME = ""
    character(len=12) :: POTSUBLIBNAME = ""
    integer :: NA(mp_MXGROUP) = 0, IPA(mp_MXGROUP + 1) = 0, PROP(mp_MXGROUP) = 1
    real(KINDDF) :: CZ(mp_MXGROUP) = 0.0d0
    real(KINDDF) :: TEMPERATURE = 0.0d0, VOLUME = 0.0d0, SPRESS0 = 0.0d0, SPRESS1 = 0.0d0
    real(KINDDF) :: SPRESS = 0.0d0, AVEPOT = 0.0d0, HARMIL = 0.0d0
    real(KINDDF) :: VTENSOR(3, 3) = 0.0d0, KTENSOR(3, 3) = 0.0d0, PTENSOR(3, 3) = 0.0d0
  end type SimMDBox

  ! --- MDRecordStamp (MD_TypeDef_RecordStamp.F90:14-31)
  type, public :: MDRecordStamp
    character(len=8) :: AppType = ""
    integer :: IProcess = 0, ITest = -1, NBox = 0, IBox(2) = -1, ITime = -1, ISect = -1
    real(KINDDF) :: Time = -1.0d0, ScalTime = -1.0d0, InstantTemp = -1.0d0
    integer :: ICfg(2) = -1, IRec(2) = -1
  end type MDRecordStamp

  ! --- SimMDCtrl (fields of MD_TypeDef_SimCtrlParam.F90:56-142 used by the force path)
  type, public :: SimMDCtrl
    real(KINDDF), allocatable :: RU(:, :), NB_RM(:, :)
    integer :: NB_UPTAB = 10, NB_MXNBS = 800, NUMFTABR = 10000, NUMFTABE = 10000
    real(KINDDF) :: RHOSCAL = 20.0d0, H = 1.0d0
    ! quench controls (MD_TypeDef_SimCtrlParam.F90:213-217): moves in lattice units RR, energy in eV
    real(KINDDF) :: STEEPEST_Alpha = 0.1d0, STEEPEST_MiStep = 1.0d-5, STEEPEST_MxStep = 0.1d0
    real(KINDDF) :: STEEPEST_MiDelE = 1.0d-3
    integer :: IFPD(3) = 1
    ! Boltzmann constant in the caller's energy unit per kelvin: CP_KB_EV for eV callers,
    ! CP_KB_CGS for the reference's CGS arrays (its CP_KB); every temperature routine reads it
    real(KINDDF) :: KB = CP_KB_EV
    ! random seeds (MD_TypeDef_SimCtrlParam.F90:51); SEED(1:2) key Thermalizing_MC_DEV
    integer :: SEED(4) = [1, 2, 3, 4]
    ! time step control (MD_TypeDef_SimCtrlParam.F90:36-47, 78-79): IHDUP < 0 searches the step
    ! from HMX every |IHDUP| steps so that no active atom moves more than DMX; the predictor damps
    ! in steps DAMPTIME0 .. DAMPTIME0 + DAMPTIME1 - 1
    integer :: IT0 = 1, IHDUP = 0, DAMPTIME0 = 1, DAMPTIME1 = 0
    real(KINDDF) :: HMI = 1.0d0, HMX = 1.0d0, DMX = 0.1d0
    ! L-BFGS(-B) quench controls (MD_TypeDef_SimCtrlParam.F90:208-210)
    real(KINDDF) :: LBFGS_PGtol = 0.0d0, LBFGS_Factr = 0.0d0
    integer :: LBFGS_Msave = 7
  end type SimMDCtrl

  ! --- MDForceClassGPU (MD_ForceClass_Register_GPU.F90:248-259): the class name selects the
  !     FS or EAM device passes; the table travels with it
  abstract interface
    ! ForceArray / EpotArray routines of AddExtForce_ForceClass (MD_ForceClass_Register_GPU.F90:
    ! 982-1180); they add their terms with AddForceFrom_Host_to_Devices / AddEpotFrom_Host_to_Devices
    subroutine EXTFORCE_PROC(SimBox, CtrlParam)
      import :: SimMDBox, SimMDCtrl
      type(SimMDBox), intent(inout) :: SimBox(:)
      type(SimMDCtrl), intent(in) :: CtrlParam
    end subroutine EXTFORCE_PROC
  end interface
  public :: EXTFORCE_PROC

  type :: ExtForceEntry
    character(len=32) :: Tag = ""
    procedure(EXTFORCE_PROC), pointer, nopass :: pForceArray => null()
    procedure(EXTFORCE_PROC), pointer, nopass :: pEpotArray => null()
  end type ExtForceEntry

  type, public :: MDForceClassGPU
    character(len=16) :: CLASSNAME = ""
    type(MDForceTable) :: ForceTable
    integer :: NEXT = 0                                  ! the ExtForceList of the reference
    type(ExtForceEntry) :: EXT(mp_MXEXTFORCE)
  end type MDForceClassGPU

  ! device context of this process (the reference keeps its device state in module variables too)
  type(c_ptr), save :: gm_SIM = c_null_ptr
  integer, save :: gm_NBOX = 0, gm_NPRT = 0
  integer, allocatable, save :: gm_NAT(:)

  interface Register_ForceTableProc
    module procedure RegisterByTypeID_ForceTableProc
    module procedure RegisterBySymbID_ForceTableProc
  end interface Register_ForceTableProc

  public :: Register_ForceTableProc, Create_Interaction_ForceTable, Register_ForceClass
  public :: SymbID2TypeID_ForceTableProc
  public :: AddExtForce_ForceClass, Clear_ExtForce, AddForceFrom_Host_to_Devices
  public :: AddEpotFrom_Host_to_Devices, CopyDISFrom_Host_to_Devices, Thermalizing_MC_DEV
  public :: Archive_Config_SimBoxArray, Restore_Config_SimBoxArray
  public :: Initialize_Globle_Variables_DEV, Init_Forcetable_Dev, Cal_NeighBoreList_DEV
  public :: CalForce_ForceClass, CalEpot_ForceClass, CalPTensor_ForceClass, Cal_AtomicStress
  public :: CalEDen_ForceClass, CopyDENFrom_Devices_to_Host
  public :: Predictor_DEV, Correction_DEV, Cal_GlobalT_DEV, VelScaling_DEV, Do_DynDamp_Forsteps_DEV
  public :: Do_Steepest_Forsteps_DEV, Do_CG_Forsteps_DEV
  public :: For_One_Step, Run_Steps_DEV, CopyOut_SimBox_DEV, Copyout_NeighboreList_DEV
  public :: CopyXPFrom_Host_to_Devices, CopyXP1From_Host_to_Devices, CheckTimestep_DEV
  public :: Autotune_DEV, Adjoint_ForceTable_DEV, Clear_Globle_Variables_DEV, MSM_Handle

contains

  subroutine msm_check(rc, where, IERR)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    integer, optional, intent(out) :: IERR
    character(len=:), allocatable :: msg
    if (present(IERR)) then
      IERR = rc
      return
    end if
    if (rc == MSM_OK) return
    msg = msm_cstring(msm_strerror(rc))
    if (c_associated(gm_SIM)) msg = msg // ": " // msm_cstring(msm_last_error(gm_SIM))
    write(*, "(A)") "MDPSCU Error: " // where // ": " // msg
    stop 1
  end subroutine msm_check

  function MSM_Handle() result(p)
    type(c_ptr) :: p
    p = gm_SIM
  end function MSM_Handle

  ! ---------------------------------------------------------------------------------------------
  ! force tables
  ! ---------------------------------------------------------------------------------------------

  ! Register_ForceTableProc, type-ID form (RegisterByTypeID_ForceTableProc,
  ! MD_TypeDef_ForceTable.F90:626-730)
  subroutine RegisterByTypeID_ForceTableProc(IFORCE, FTable, NNFORCE, NEFORCE, EMBDF, NOTE)
    integer, intent(in) :: IFORCE
    type(MDForceTable), intent(inout) :: FTable
    procedure(NNFORCE_PROC) :: NNFORCE
    procedure(NEFORCE_PROC), optional :: NEFORCE
    procedure(EMBDF_PROC), optional :: EMBDF
    character(len=*), optional, intent(in) :: NOTE
    if (IFORCE < 1 .or. IFORCE > mp_MXFORCE) then
      write(*, "(A,I0)") "MDPSCU Error: potential type id out of range: ", IFORCE
      stop 1
    end if
    FTable%PROC(IFORCE)%REG = .true.
    FTable%PROC(IFORCE)%pNN => NNFORCE
    if (present(NEFORCE)) FTable%PROC(IFORCE)%pNE => NEFORCE
    if (present(EMBDF)) FTable%PROC(IFORCE)%pEMB => EMBDF
    if (present(NOTE)) FTable%PROC(IFORCE)%NOTE = NOTE
  end subroutine RegisterByTypeID_ForceTableProc

  function upcase8(str) result(u)
    character(len=*), intent(in) :: str
    character(len=8) :: u
    integer :: i, c
    u = str
    do i = 1, len_trim(u)
      c = iachar(u(i:i))
      if (c >= iachar("a") .and. c <= iachar("z")) u(i:i) = achar(c - 32)
    end do
  end function upcase8

  ! Register_ForceTableProc, symbol form (RegisterBySymbID_ForceTableProc,
  ! MD_TypeDef_ForceTable.F90:734-852): the routines of the interaction of target atoms Tgt with
  ! source atoms Sor, symbol "TGT<-SOR" (8 characters, upper case); the type id is one more than
  ! the largest registered; a symbol registered twice stops
  subroutine RegisterBySymbID_ForceTableProc(Sor, Tgt, FTable, NNFORCE, NEFORCE, EMBDF, NOTE)
    character(len=*), intent(in) :: Sor, Tgt
    type(MDForceTable), intent(inout) :: FTable
    procedure(NNFORCE_PROC) :: NNFORCE
    procedure(NEFORCE_PROC), optional :: NEFORCE
    procedure(EMBDF_PROC), optional :: EMBDF
    character(len=*), optional, intent(in) :: NOTE
    character(len=8) :: symb
    integer :: i, id
    symb = upcase8(trim(Tgt) // "<-" // trim(Sor))
    id = 0
    do i = 1, mp_MXFORCE
      if (.not. FTable%PROC(i)%REG) cycle
      if (FTable%PROC(i)%SYMB == symb) then
        write(*, "(A)") "MDPSCU Error: the potential type by symb" // trim(symb) // &
                        " has been registed"
        stop 1
      end if
      id = i
    end do
    id = id + 1
    call RegisterByTypeID_ForceTableProc(id, FTable, NNFORCE, NEFORCE, EMBDF, NOTE)
    FTable%PROC(id)%SYMB = symb
  end subroutine RegisterBySymbID_ForceTableProc

  ! SymbID2TypeID_ForceTableProc (MD_TypeDef_ForceTable.F90:856-886): type id registered for
  ! "TGT<-SOR", 0 if none
  integer function SymbID2TypeID_ForceTableProc(Sor, Tgt, FTable) result(IFORCE)
    character(len=*), intent(in) :: Sor, Tgt
    type(MDForceTable), intent(in) :: FTable
    character(len=8) :: symb
    integer :: i
    symb = upcase8(trim(Tgt) // "<-" // trim(Sor))
    IFORCE = 0
    do i = 1, mp_MXFORCE
      if (FTable%PROC(i)%REG .and. FTable%PROC(i)%SYMB == symb) then
        IFORCE = i
        return
      end if
    end do
  end function SymbID2TypeID_ForceTableProc

  integer function find_or_add(list, n, v) result(k)
    integer, intent(inout) :: list(:)
    integer, intent(inout) :: n
    integer, intent(in) :: v
    do k = 1, n
      if (list(k) == v) return
    end do
    n = n + 1
    list(n) = v
    k = n
  end function find_or_add

  ! Create_Interaction_ForceTable (MD_TypeDef_ForceTable.F90:1151-1230) with New_ForceTable
  ! (:530-622), Create_Pairwise_ForceTable (:890-982) and Create_EMBDFUNTable (:986-1056):
  ! distinct pair types of SimBox%PTYPE become table kinds; knot i sits at r = (i/CSI)^2 with
  ! CSI = NTAB/sqrt(Rmax), Rmax = maxval(RU); POTR = NN*r, FPOTR = dNN*r, POTB/FPOTB = NE;
  ! RHOMX = RHOSCAL * max POTB, RHOD = RHOMX/NEMBD, F tabulated at rho = (i-1)*RHOD.
  subroutine Create_Interaction_ForceTable(SimBox, CtrlParam, FTable)
    type(SimMDBox), intent(inout) :: SimBox
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceTable), intent(inout) :: FTable
    integer :: ng, i, j, k, n1, n2, it
    integer :: tmp(mp_MXFORCE), tmp1(mp_MXFORCE)
    real(KINDDF) :: r, t, v, dv
    ng = SimBox%NGROUP
    n1 = 0
    n2 = 0
    if (allocated(FTable%KPAIR)) deallocate(FTable%KPAIR)
    allocate(FTable%KPAIR(ng, ng))
    ! pairs without a type id take the one registered for "SYMB(i)<-SYMB(j)" (:1170-1185)
    do j = 1, ng
      do i = 1, ng
        if (SimBox%PTYPE(i, j) == 0 .and. len_trim(SimBox%SYMB(i)) > 0) &
          SimBox%PTYPE(i, j) = SymbID2TypeID_ForceTableProc(SimBox%SYMB(j), SimBox%SYMB(i), FTable)
      end do
    end do
    do j = 1, ng
      do i = 1, ng
        it = SimBox%PTYPE(i, j)
        if (it < 1 .or. it > mp_MXFORCE) then
          write(*, "(A,I0,A,I0)") "MDPSCU Error: cannot find the registed force table for " // &
                                  "between atom groups ", i, " and ", j
          stop 1
        end if
        if (.not. FTable%PROC(it)%REG) then
          write(*, "(A,I0,A)") "MDPSCU Error: the potential type #", it, " is not available"
          stop 1
        end if
        FTable%KPAIR(i, j) = find_or_add(tmp, n1, it)
      end do
    end do
    FTable%NKIND = n1
    FTable%NTAB = CtrlParam%NUMFTABR
    FTable%Rmax = maxval(CtrlParam%RU)
    FTable%RmaxSQRT = sqrt(FTable%Rmax)
    FTable%CSI = dble(FTable%NTAB) / FTable%RmaxSQRT
    FTable%CSIV = 1.0d0 / FTable%CSI
    FTable%RHOMX = -1.0d0
    FTable%RHOMULT = CtrlParam%RHOSCAL
    if (allocated(FTable%FPAIR)) deallocate(FTable%FPAIR)
    allocate(FTable%FPAIR(n1))
    FTable%FPAIR = tmp(1:n1)
    if (allocated(FTable%POTR)) deallocate(FTable%POTR, FTable%FPOTR, FTable%POTB, FTable%FPOTB)
    allocate(FTable%POTR(n1, FTable%NTAB), FTable%FPOTR(n1, FTable%NTAB), &
             FTable%POTB(n1, FTable%NTAB), FTable%FPOTB(n1, FTable%NTAB))
    do k = 1, n1
      it = FTable%FPAIR(k)
      do i = 1, FTable%NTAB
        t = dble(i) * FTable%CSIV
        r = t * t
        call FTable%PROC(it)%pNN(r, v, dv)
        FTable%POTR(k, i) = v * r
        FTable%FPOTR(k, i) = dv * r
        if (associated(FTable%PROC(it)%pNE)) then
          call FTable%PROC(it)%pNE(r, FTable%POTB(k, i), FTable%FPOTB(k, i))
          if (FTable%RHOMX < FTable%POTB(k, i) * FTable%RHOMULT) &
            FTable%RHOMX = FTable%POTB(k, i) * FTable%RHOMULT
        else
          FTable%POTB(k, i) = 0.0d0
          FTable%FPOTB(k, i) = 0.0d0
        end if
      end do
    end do
    FTable%NEMBD = CtrlParam%NUMFTABE
    FTable%RHOD = FTable%RHOMX / dble(FTable%NEMBD)
    if (trim(FTable%PotType) == "EAM_TYPE") then
      if (allocated(FTable%KEMBD)) deallocate(FTable%KEMBD)
      allocate(FTable%KEMBD(ng))
      do i = 1, ng
        it = SimBox%PTYPE(i, i)
        if (.not. associated(FTable%PROC(it)%pEMB)) then
          write(*, "(A,I0,A)") "MDPSCU Error: the embedment function of potential type #", it, &
                               " is not available"
          stop 1
        end if
        FTable%KEMBD(i) = find_or_add(tmp1, n2, it)
      end do
      FTable%NKIND1 = n2
      if (allocated(FTable%FPAIR1)) deallocate(FTable%FPAIR1)
      allocate(FTable%FPAIR1(n2))
      FTable%FPAIR1 = tmp1(1:n2)
      if (FTable%RHOD <= 1.0d-64) FTable%RHOD = 1.0d0
      if (allocated(FTable%FEMBD)) deallocate(FTable%FEMBD, FTable%DFEMBD)
      allocate(FTable%FEMBD(n2, FTable%NEMBD), FTable%DFEMBD(n2, FTable%NEMBD))
      do k = 1, n2
        it = FTable%FPAIR1(k)
        do i = 1, FTable%NEMBD
          r = dble(i - 1) * FTable%RHOD
          call FTable%PROC(it)%pEMB(r, FTable%FEMBD(k, i), FTable%DFEMBD(k, i))
        end do
      end do
    end if
  end subroutine Create_Interaction_ForceTable

  ! Register_ForceClass (MD_ForceClass_Register_GPU.F90:363-442): FS_TYPE or EAM_TYPE
  subroutine Register_ForceClass(classname, ForceClass)
    character(len=*), intent(in) :: classname
    type(MDForceClassGPU), intent(inout) :: ForceClass
    select case (trim(adjustl(classname)))
    case ("FS_TYPE", "EAM_TYPE")
      ForceClass%CLASSNAME = trim(adjustl(classname))
      ForceClass%ForceTable%PotType = ForceClass%CLASSNAME
    case default
      write(*, "(A)") "MDPSCU Error: unsupported type of potential " // trim(classname)
      stop 1
    end select
  end subroutine Register_ForceClass

  ! ---------------------------------------------------------------------------------------------
  ! device context
  ! ---------------------------------------------------------------------------------------------

  ! Initialize_Globle_Variables_DEV (MD_Globle_Variables_GPU.F90:515-654), generalised to boxes
  ! of uneven size: context, control data, boxes and atoms (concatenated in box order).
  subroutine Initialize_Globle_Variables_DEV(SimBox, CtrlParam, DEVICE, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    integer, optional, intent(in) :: DEVICE
    integer, optional, intent(out) :: IERR
    integer :: nb, b, n, off, ng, dev
    integer(c_int), allocatable :: nat(:), pd(:), ity(:), st(:)
    real(c_double), allocatable :: lo(:), ln(:), xp(:), xp1(:)
    integer(c_int) :: rc
    dev = 0
    if (present(DEVICE)) dev = DEVICE
    if (.not. c_associated(gm_SIM)) then
      rc = msm_create(gm_SIM, int(dev, c_int))
      call msm_check(rc, "Initialize_Globle_Variables_DEV", IERR)
      if (rc /= MSM_OK) return
    end if
    nb = size(SimBox)
    ng = SimBox(1)%NGROUP
    rc = msm_set_ctrl(gm_SIM, int(ng, c_int), CtrlParam%RU, CtrlParam%NB_RM, &
                      int(CtrlParam%NB_MXNBS, c_int), SimBox(1)%CM)
    call msm_check(rc, "Initialize_Globle_Variables_DEV", IERR)
    if (rc /= MSM_OK) return
    n = sum(SimBox(:)%NPRT)
    allocate(nat(nb), pd(3 * nb), lo(3 * nb), ln(3 * nb))
    allocate(xp(3 * n), xp1(3 * n), ity(n), st(n))
    off = 0
    do b = 1, nb
      nat(b) = SimBox(b)%NPRT
      lo(3 * b - 2:3 * b) = SimBox(b)%BOXLOW
      ln(3 * b - 2:3 * b) = SimBox(b)%ZL
      pd(3 * b - 2:3 * b) = CtrlParam%IFPD
    end do
    ! XP(N,3) of the concatenation: component-major over all atoms
    do b = 1, nb
      xp(off + 1:off + SimBox(b)%NPRT) = SimBox(b)%XP(:, 1)
      xp(n + off + 1:n + off + SimBox(b)%NPRT) = SimBox(b)%XP(:, 2)
      xp(2 * n + off + 1:2 * n + off + SimBox(b)%NPRT) = SimBox(b)%XP(:, 3)
      xp1(off + 1:off + SimBox(b)%NPRT) = SimBox(b)%XP1(:, 1)
      xp1(n + off + 1:n + off + SimBox(b)%NPRT) = SimBox(b)%XP1(:, 2)
      xp1(2 * n + off + 1:2 * n + off + SimBox(b)%NPRT) = SimBox(b)%XP1(:, 3)
      ity(off + 1:off + SimBox(b)%NPRT) = SimBox(b)%ITYP
      st(off + 1:off + SimBox(b)%NPRT) = SimBox(b)%STATU
      off = off + SimBox(b)%NPRT
    end do
    rc = msm_set_boxes(gm_SIM, int(nb, c_int), nat, lo, ln, pd)
    call msm_check(rc, "Initialize_Globle_Variables_DEV", IERR)
    if (rc /= MSM_OK) return
    rc = msm_set_atoms(gm_SIM, int(n, c_int), xp, xp1, ity, st)
    call msm_check(rc, "Initialize_Globle_Variables_DEV", IERR)
    gm_NBOX = nb
    gm_NPRT = n
    if (allocated(gm_NAT)) deallocate(gm_NAT)
    allocate(gm_NAT(nb))
    gm_NAT = nat
  end subroutine Initialize_Globle_Variables_DEV

  ! Init_Forcetable_Dev / INITIALIZE_FS_Force_Table_DEV / INITIALIZE_EAM_Force_Table_DEV
  subroutine Init_Forcetable_Dev(SimBox, CtrlParam, ForceClass, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(inout), target :: ForceClass
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc, pt, ng
    integer(c_int), allocatable, target :: kembd(:)
    type(c_ptr) :: pk, pf, pd
    associate(T => ForceClass%ForceTable)
      ng = SimBox(1)%NGROUP
      pt = MSM_FS_TYPE
      pk = c_null_ptr
      pf = c_null_ptr
      pd = c_null_ptr
      if (trim(T%PotType) == "EAM_TYPE") then
        pt = MSM_EAM_TYPE
        allocate(kembd(ng))
        kembd = T%KEMBD
        pk = c_loc(kembd)
        pf = c_loc(T%FEMBD)
        pd = c_loc(T%DFEMBD)
      end if
      rc = msm_set_force_table(gm_SIM, pt, ng, int(T%NKIND, c_int), int(T%NTAB, c_int), T%CSI, &
                               T%KPAIR, T%POTR, T%FPOTR, T%POTB, T%FPOTB, &
                               int(T%NKIND1, c_int), int(T%NEMBD, c_int), T%RHOD, pk, pf, pd)
    end associate
    call msm_check(rc, "Init_Forcetable_Dev", IERR)
  end subroutine Init_Forcetable_Dev

  ! Cal_NeighBoreList_DEV (MD_NeighborsList_GPU.F90:121-181)
  subroutine Cal_NeighBoreList_DEV(SimBox, CtrlParam, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    integer, optional, intent(out) :: IERR
    call msm_check(msm_build_neighbors(gm_SIM), "Cal_NeighBoreList_DEV", IERR)
  end subroutine Cal_NeighBoreList_DEV

  ! CalForce_ForceClass (MD_ForceClass_Register_GPU.F90:636-680): density then force pass
  ! then Cumulate_ExtForce (:584-632): the registered ForceArray routines in registration order
  subroutine CalForce_ForceClass(SimBox, CtrlParam, ForceClass, IERR)
    type(SimMDBox), intent(inout) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc
    integer :: k
    rc = msm_calc_force(gm_SIM)
    call msm_check(rc, "CalForce_ForceClass", IERR)
    if (rc /= MSM_OK) return
    do k = 1, ForceClass%NEXT
      if (associated(ForceClass%EXT(k)%pForceArray)) &
        call ForceClass%EXT(k)%pForceArray(SimBox, CtrlParam)
    end do
  end subroutine CalForce_ForceClass

  ! the pCalEDen slot of MDForceClassGPU (MD_ForceClass_Register_GPU.F90:254, 405, 416):
  ! CALDEN_FS/EAM_Force_Table2A_DEV (MD_FS_ForceTable_GPU.F90:875-930), the density pass alone;
  ! F'(rho) stays on the device (CopyDENFrom_Devices_to_Host copies it out)
  subroutine CalEDen_ForceClass(SimBox, CtrlParam, ForceClass, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    integer, optional, intent(out) :: IERR
    call msm_check(msm_calc_density(gm_SIM), "CalEDen_ForceClass", IERR)
  end subroutine CalEDen_ForceClass

  ! the device array DEN (F'(rho) per atom) in the caller's atom order of all boxes
  subroutine CopyDENFrom_Devices_to_Host(DEN, IERR)
    real(KINDDF), intent(out) :: DEN(:)
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc
    if (size(DEN) < gm_NPRT) then
      rc = MSM_ERR_ARG
    else
      rc = msm_get_den(gm_SIM, DEN)
    end if
    call msm_check(rc, "CopyDENFrom_Devices_to_Host", IERR)
  end subroutine CopyDENFrom_Devices_to_Host

  ! CalEpot_ForceClass (:778-822)
  subroutine CalEpot_ForceClass(SimBox, CtrlParam, ForceClass, IERR)
    type(SimMDBox), intent(inout) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc
    integer :: k
    rc = msm_calc_epot(gm_SIM)
    call msm_check(rc, "CalEpot_ForceClass", IERR)
    if (rc /= MSM_OK) return
    do k = 1, ForceClass%NEXT
      if (associated(ForceClass%EXT(k)%pEpotArray)) &
        call ForceClass%EXT(k)%pEpotArray(SimBox, CtrlParam)
    end do
  end subroutine CalEpot_ForceClass

  ! CalPTensor_ForceClass (MD_ForceClass_Register_GPU.F90:684-728): the virial tensor as
  ! COPYOUT_VIRIALTENSOR returns it (MD_FS_ForceTable_GPU.F90:1415-1447), the mean over the boxes
  subroutine CalPTensor_ForceClass(SimBox, CtrlParam, ForceClass, VTENSOR, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    real(KINDDF), intent(out) :: VTENSOR(3, 3)
    integer, optional, intent(out) :: IERR
    real(c_double), allocatable :: v(:)
    integer(c_int) :: rc
    integer :: b
    allocate(v(9 * gm_NBOX))
    rc = msm_calc_ptensor(gm_SIM, v)
    call msm_check(rc, "CalPTensor_ForceClass", IERR)
    VTENSOR = 0.0d0
    if (rc /= MSM_OK) return
    do b = 1, gm_NBOX
      VTENSOR = VTENSOR + reshape(v(9 * b - 8:9 * b), [3, 3])
    end do
    VTENSOR = VTENSOR / dble(gm_NBOX)
  end subroutine CalPTensor_ForceClass

  ! Cal_AtomicStress (Analysis/AtomicVirialStress_GPU.F90:734-773): the virial part of the atomic
  ! pressure tensor through CAL_FS/EAM_AtomicStress_KERNEL, hAVP(NPRT, 9) over all boxes in the
  ! caller's atom order, components P11, P12, P13, P21, ..., P33 (zero for inactive atoms)
  subroutine Cal_AtomicStress(SimBox, CtrlParam, hAVP, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    real(KINDDF), intent(out) :: hAVP(:, :)
    integer, optional, intent(out) :: IERR
    real(c_double), allocatable :: v(:)
    integer(c_int) :: rc
    hAVP = 0.0d0
    if (size(hAVP, 1) < gm_NPRT .or. size(hAVP, 2) < 9) then
      if (present(IERR)) then
        IERR = MSM_ERR_ARG
        return
      end if
      write(*, "(A)") "MDPSCU Error: Cal_AtomicStress: hAVP must be at least (NPRT, 9)"
      stop 1
    end if
    allocate(v(9 * max(gm_NPRT, 1)))
    rc = msm_calc_atomic_stress(gm_SIM, v)
    call msm_check(rc, "Cal_AtomicStress", IERR)
    if (rc /= MSM_OK) return
    hAVP(1:gm_NPRT, 1:9) = reshape(v(1:9 * gm_NPRT), [gm_NPRT, 9])
  end subroutine Cal_AtomicStress

  ! Cal_GlobalT_DEV (MD_DiffScheme_GPU.F90:1042-1064): 2 <EK> / (3 kB) over the atoms of all boxes
  ! that count (active, no FIXPOS bit), in K
  subroutine Cal_GlobalT_DEV(SimBox, CtrlParam, CURT, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    real(KINDDF), intent(out) :: CURT
    integer, optional, intent(out) :: IERR
    real(c_double), allocatable :: ek(:), tb(:)
    integer(c_int) :: rc
    CURT = 0.0d0
    allocate(ek(max(gm_NPRT, 1)), tb(max(gm_NBOX, 1)))
    rc = msm_calc_ekin(gm_SIM, CtrlParam%KB, ek, tb)
    call msm_check(rc, "Cal_GlobalT_DEV", IERR)
    if (rc /= MSM_OK) return
    if (count(ek(1:gm_NPRT) >= 0.0d0) > 0) &
      CURT = 2.0d0 * sum(ek(1:gm_NPRT), mask=(ek(1:gm_NPRT) >= 0.0d0)) / &
             dble(count(ek(1:gm_NPRT) >= 0.0d0)) / (3.0d0 * CtrlParam%KB)
  end subroutine Cal_GlobalT_DEV

  ! VelScaling_DEV (MD_DiffScheme_GPU.F90:1390-1446): scale the velocities of each box to DT (K)
  subroutine VelScaling_DEV(SimBox, CtrlParam, DT, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    real(KINDDF), intent(in) :: DT
    integer, optional, intent(out) :: IERR
    call msm_check(msm_scale_velocities(gm_SIM, DT, CtrlParam%KB), "VelScaling_DEV", IERR)
  end subroutine VelScaling_DEV

  ! Do_DynDamp_Forsteps_DEV (MD_DiffScheme_GPU.F90:1809-1883): damped dynamics on the current
  ! neighbour list until the largest per-atom change of EPOT is at most STEEPEST_MiDelE (eV)
  subroutine Do_DynDamp_Forsteps_DEV(SimBox, CtrlParam, ForceClass, MXNUMSTEPS, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    integer, intent(in) :: MXNUMSTEPS
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc, iter
    real(c_double) :: delepot
    rc = msm_dyn_damp(gm_SIM, MXNUMSTEPS, CtrlParam%STEEPEST_MiDelE, CtrlParam%H, iter, delepot)
    call msm_check(rc, "Do_DynDamp_Forsteps_DEV", IERR)
    if (rc /= MSM_OK) return
    if (iter > MXNUMSTEPS) then
      write(*, "(A, I8)") " MDPSCU WARNING: DnyDamp finished after max steps:  ", iter - 1
    else
      write(*, "(A, I8)") " MDPSCU Message: DnyDamp finished after steps:  ", iter
    end if
    write(*, "(A, 1PE13.4)") "                 with max energy uncertainty(ev): ", delepot
  end subroutine Do_DynDamp_Forsteps_DEV

  ! Do_Steepest_Forsteps_DEV (MD_SteepestScheme_GPU.F90:264-290): steepest descent, with the line
  ! search when METH has the CP_DAMPSCHEME_LSEARCH bit (energy tolerance 1e-3 eV, as the
  ! reference); moves bounded by STEEPEST_MxStep / STEEPEST_MiStep times SimBox(1)%RR
  subroutine Do_Steepest_Forsteps_DEV(SimBox, CtrlParam, ForceClass, MXNUMSTEPS, METH, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    integer, intent(in) :: MXNUMSTEPS, METH
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc, iter, ls
    real(c_double) :: delepot, maxmove, mindele
    ls = 0
    mindele = CtrlParam%STEEPEST_MiDelE
    if (iand(METH, CP_DAMPSCHEME_LSEARCH) > 0) then
      ls = 1
      mindele = 1.0d-3
    end if
    rc = msm_steepest(gm_SIM, MXNUMSTEPS, ls, CtrlParam%STEEPEST_Alpha, &
                      CtrlParam%STEEPEST_MxStep * SimBox(1)%RR, &
                      CtrlParam%STEEPEST_MiStep * SimBox(1)%RR, mindele, iter, delepot, maxmove)
    call msm_check(rc, "Do_Steepest_Forsteps_DEV", IERR)
    if (rc /= MSM_OK) return
    if (iter > MXNUMSTEPS) then
      write(*, "(A, I8)") " MDPSCU WARNING: steepest finished after max steps: ", iter
    else
      write(*, "(A, I8)") " MDPSCU Message: steepest finished after steps: ", iter
    end if
    write(*, "(A, 1PE13.4)") "                 with max energy uncertainty(ev): ", delepot
  end subroutine Do_Steepest_Forsteps_DEV

  ! Do_CG_Forsteps_DEV (MD_CGScheme_GPU.F90:296-322): conjugate gradients (Polak-Ribiere), with
  ! secant line searches when METH has the CP_DAMPSCHEME_LSEARCH bit
  subroutine Do_CG_Forsteps_DEV(SimBox, CtrlParam, ForceClass, MXNUMSTEPS, METH, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    integer, intent(in) :: MXNUMSTEPS, METH
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc, iter, ls
    real(c_double) :: delepot
    ls = 0
    if (iand(METH, CP_DAMPSCHEME_LSEARCH) > 0) ls = 1
    rc = msm_cg(gm_SIM, MXNUMSTEPS, ls, CtrlParam%STEEPEST_MxStep * SimBox(1)%RR, &
                CtrlParam%STEEPEST_MiStep * SimBox(1)%RR, CtrlParam%STEEPEST_MiDelE, iter, delepot)
    call msm_check(rc, "Do_CG_Forsteps_DEV", IERR)
    if (rc /= MSM_OK) return
    if (iter > MXNUMSTEPS) then
      write(*, "(A, I8)") " MDPSCU WARNING: CG finished after max steps: ", iter - 1
    else
      write(*, "(A, I8)") " MDPSCU Message: CG finished after steps: ", iter
    end if
    write(*, "(A, 1PE13.4)") "                 with max energy uncertainty(ev): ", delepot
  end subroutine Do_CG_Forsteps_DEV

  ! CheckTimestep_DEV (MD_DiffScheme_GPU.F90:1068-1210): IFLAG = 1 when some active atom would
  ! move more than sqrt(MXD2) in a step TH (the variable-step search halves TH until IFLAG = 0)
  subroutine CheckTimestep_DEV(TH, MXD2, IFLAG, IERR)
    real(c_double), intent(in) :: TH, MXD2
    integer, intent(out) :: IFLAG
    integer, optional, intent(out) :: IERR
    integer(c_int) :: fl
    fl = 0
    call msm_check(msm_check_timestep(gm_SIM, TH, MXD2, fl), "CheckTimestep_DEV", IERR)
    IFLAG = fl
  end subroutine CheckTimestep_DEV

  ! Predictor_DEV (MD_DiffScheme_GPU.F90:604-682): damping (DAMPING_template) when ITIME lies in
  ! DAMPTIME0 .. DAMPTIME0 + DAMPTIME1 - 1; with IHDUP < 0 and mod(ITIME - IT0 + 1, |IHDUP|) = 0
  ! the step search of :629-657 -- from HMX, halve the step while CheckTimestep_DEV finds an
  ! active atom moving more than DMX -- which rewrites CtrlParam%H; then the Swope predictor
  subroutine Predictor_DEV(ITIME, SimBox, CtrlParam, IERR)
    integer, intent(in) :: ITIME
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(inout) :: CtrlParam
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc, fl
    real(c_double) :: th
    rc = MSM_OK
    if (ITIME >= CtrlParam%DAMPTIME0 .and. &
        ITIME <= CtrlParam%DAMPTIME0 + CtrlParam%DAMPTIME1 - 1) rc = msm_damping(gm_SIM)
    if (rc == MSM_OK .and. CtrlParam%IHDUP < 0) then
      if (mod(ITIME - CtrlParam%IT0 + 1, abs(CtrlParam%IHDUP)) == 0) then
        th = CtrlParam%HMX
        do
          fl = 0
          rc = msm_check_timestep(gm_SIM, th, CtrlParam%DMX * CtrlParam%DMX, fl)
          if (rc /= MSM_OK .or. fl == 0) exit
          th = 0.5d0 * th
          if (th == 0.0d0) then
            rc = MSM_ERR_ARG
            exit
          end if
        end do
        if (rc == MSM_OK) CtrlParam%H = th
      end if
    end if
    if (rc == MSM_OK) rc = msm_predictor(gm_SIM, CtrlParam%H)
    call msm_check(rc, "Predictor_DEV", IERR)
  end subroutine Predictor_DEV

  subroutine Correction_DEV(ITIME, SimBox, CtrlParam, IERR)
    integer, intent(in) :: ITIME
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    integer, optional, intent(out) :: IERR
    call msm_check(msm_correction(gm_SIM, CtrlParam%H), "Correction_DEV", IERR)
  end subroutine Correction_DEV

  ! For_One_Step (MD_Method_GenericMD_GPU.F90:496-659): predictor, list rebuild when
  ! mod(ITIME, NB_UPTAB) = 0, force, corrector -- one launch sequence per call
  ! (damping window, step search and external forces included, as Predictor_DEV and
  ! CalForce_ForceClass do them)
  subroutine For_One_Step(ITIME, SimBox, CtrlParam, ForceClass, IERR)
    integer, intent(in) :: ITIME
    type(SimMDBox), intent(inout) :: SimBox(:)
    type(SimMDCtrl), intent(inout) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    integer, optional, intent(out) :: IERR
    integer :: ie
    integer(c_int) :: rc
    call Predictor_DEV(ITIME, SimBox, CtrlParam, ie)
    rc = ie
    if (rc == MSM_OK .and. CtrlParam%NB_UPTAB > 0) then
      if (mod(ITIME, CtrlParam%NB_UPTAB) == 0) rc = msm_build_neighbors(gm_SIM)
    end if
    if (rc == MSM_OK) then
      call CalForce_ForceClass(SimBox, CtrlParam, ForceClass, ie)
      rc = ie
    end if
    if (rc == MSM_OK) rc = msm_correction(gm_SIM, CtrlParam%H)
    call msm_check(rc, "For_One_Step", IERR)
  end subroutine For_One_Step

  ! NSTEPS steps ITIME0+1 .. ITIME0+NSTEPS of For_One_Step semantics in one call (fused kernels,
  ! replayed as captured graphs); forces of the current positions must be current on entry.
  subroutine Run_Steps_DEV(ITIME0, NSTEPS, SimBox, CtrlParam, IERR)
    integer, intent(in) :: ITIME0, NSTEPS
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    integer, optional, intent(out) :: IERR
    call msm_check(msm_run(gm_SIM, int(NSTEPS, c_int), int(CtrlParam%NB_UPTAB, c_int), &
                           int(ITIME0, c_int), CtrlParam%H), "Run_Steps_DEV", IERR)
  end subroutine Run_Steps_DEV

  ! CopyOut_SimBox_DEV: XP, XP1, FP, DIS, STATU (and EPOT) back into the boxes
  subroutine CopyOut_SimBox_DEV(SimBox, WITHEPOT, IERR)
    type(SimMDBox), intent(inout) :: SimBox(:)
    logical, optional, intent(in) :: WITHEPOT
    integer, optional, intent(out) :: IERR
    real(c_double), allocatable, target :: xp(:), xp1(:), fp(:), dis(:), ep(:)
    integer(c_int), allocatable, target :: st(:)
    integer(c_int) :: rc
    integer :: n, b, off, k, m
    n = gm_NPRT
    allocate(xp(3 * n), xp1(3 * n), fp(3 * n), dis(3 * n), st(n), ep(n))
    rc = msm_get_atoms(gm_SIM, c_loc(xp), c_loc(xp1), c_loc(fp), c_loc(st), c_loc(dis))
    if (rc == MSM_OK .and. present(WITHEPOT)) then
      if (WITHEPOT) rc = msm_get_epot(gm_SIM, ep)
    end if
    call msm_check(rc, "CopyOut_SimBox_DEV", IERR)
    if (rc /= MSM_OK) return
    off = 0
    do b = 1, size(SimBox)
      m = SimBox(b)%NPRT
      if (.not. allocated(SimBox(b)%FP)) allocate(SimBox(b)%FP(m, 3))
      if (.not. allocated(SimBox(b)%DIS)) allocate(SimBox(b)%DIS(m, 3))
      if (.not. allocated(SimBox(b)%EPOT)) allocate(SimBox(b)%EPOT(m))
      do k = 1, 3
        SimBox(b)%XP(:, k) = xp((k - 1) * n + off + 1:(k - 1) * n + off + m)
        SimBox(b)%XP1(:, k) = xp1((k - 1) * n + off + 1:(k - 1) * n + off + m)
        SimBox(b)%FP(:, k) = fp((k - 1) * n + off + 1:(k - 1) * n + off + m)
        SimBox(b)%DIS(:, k) = dis((k - 1) * n + off + 1:(k - 1) * n + off + m)
      end do
      SimBox(b)%STATU = st(off + 1:off + m)
      if (present(WITHEPOT)) then
        if (WITHEPOT) SimBox(b)%EPOT = ep(off + 1:off + m)
      end if
      off = off + m
    end do
  end subroutine CopyOut_SimBox_DEV

  ! CopyXPFrom_Host_to_Devices / CopyXP1From_Host_to_Devices (MD_Globle_Variables_GPU.F90:349-373):
  ! SimBox positions or velocities of all boxes onto the device, no list rebuild -- what the
  ! host-side quench loops (the L-BFGS driver) use between force evaluations
  subroutine CopyXPFrom_Host_to_Devices(SimBox, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    integer, optional, intent(out) :: IERR
    call put_xyz(SimBox, .false., "CopyXPFrom_Host_to_Devices", IERR)
  end subroutine CopyXPFrom_Host_to_Devices

  subroutine CopyXP1From_Host_to_Devices(SimBox, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    integer, optional, intent(out) :: IERR
    call put_xyz(SimBox, .true., "CopyXP1From_Host_to_Devices", IERR)
  end subroutine CopyXP1From_Host_to_Devices

  subroutine put_xyz(SimBox, VEL, WHERE, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    logical, intent(in) :: VEL
    character(*), intent(in) :: WHERE
    integer, optional, intent(out) :: IERR
    real(c_double), allocatable, target :: a(:)
    integer :: n, b, off, k, m
    integer(c_int) :: rc
    n = gm_NPRT
    allocate(a(3 * n))
    off = 0
    do b = 1, size(SimBox)
      m = SimBox(b)%NPRT
      do k = 1, 3
        if (VEL) then
          a((k - 1) * n + off + 1:(k - 1) * n + off + m) = SimBox(b)%XP1(:, k)
        else
          a((k - 1) * n + off + 1:(k - 1) * n + off + m) = SimBox(b)%XP(:, k)
        end if
      end do
      off = off + m
    end do
    if (off /= n) then
      rc = MSM_ERR_ARG
    else if (VEL) then
      rc = msm_put_atoms(gm_SIM, c_null_ptr, c_loc(a))
    else
      rc = msm_put_atoms(gm_SIM, c_loc(a), c_null_ptr)
    end if
    call msm_check(rc, WHERE, IERR)
  end subroutine put_xyz

  ! Copyout_NeighboreList_DEV (MD_NeighborsList_GPU.F90:604-693): the full list in the caller's
  ! atom order, 1-based, INDI(N, MXKVOIS)
  subroutine Copyout_NeighboreList_DEV(KVOIS, INDI, IERR)
    integer(c_int), intent(out) :: KVOIS(:), INDI(:, :)
    integer, optional, intent(out) :: IERR
    call msm_check(msm_get_neighbors(gm_SIM, KVOIS, INDI, int(size(INDI, 1), c_int)), &
                   "Copyout_NeighboreList_DEV", IERR)
  end subroutine Copyout_NeighboreList_DEV

  ! Kernel-variant autotuner on the current state (msm_autotune)
  subroutine Autotune_DEV(REPS, CHOICE, IERR)
    integer, intent(in) :: REPS
    integer(c_int), intent(out) :: CHOICE(8)
    integer, optional, intent(out) :: IERR
    real(c_double) :: ms(2)
    call msm_check(msm_autotune(gm_SIM, int(REPS, c_int), CHOICE, ms), "Autotune_DEV", IERR)
  end subroutine Autotune_DEV

  ! Table-knot gradients of a loss L(FP, EPOT per box) (msm_adjoint): DLDF(N,3) in the
  ! concatenated atom order, DLDE(NBOX); GRAD receives dL/d(knot) in the table layout of FTable
  ! (one table kind).
  subroutine Adjoint_ForceTable_DEV(DLDF, DLDE, FTable, GRAD, IERR)
    real(KINDDF), intent(in) :: DLDF(:, :), DLDE(:)
    type(MDForceTable), intent(in) :: FTable
    type(MDForceTable), intent(inout), target :: GRAD
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc
    type(c_ptr) :: pf, pd
    GRAD%PotType = FTable%PotType
    GRAD%NKIND = FTable%NKIND
    GRAD%NTAB = FTable%NTAB
    GRAD%NKIND1 = FTable%NKIND1
    GRAD%NEMBD = FTable%NEMBD
    if (allocated(GRAD%POTR)) deallocate(GRAD%POTR, GRAD%FPOTR, GRAD%POTB, GRAD%FPOTB)
    allocate(GRAD%POTR(FTable%NKIND, FTable%NTAB), GRAD%FPOTR(FTable%NKIND, FTable%NTAB), &
             GRAD%POTB(FTable%NKIND, FTable%NTAB), GRAD%FPOTB(FTable%NKIND, FTable%NTAB))
    pf = c_null_ptr
    pd = c_null_ptr
    if (trim(FTable%PotType) == "EAM_TYPE") then
      if (allocated(GRAD%FEMBD)) deallocate(GRAD%FEMBD, GRAD%DFEMBD)
      allocate(GRAD%FEMBD(FTable%NKIND1, FTable%NEMBD), GRAD%DFEMBD(FTable%NKIND1, FTable%NEMBD))
      pf = c_loc(GRAD%FEMBD)
      pd = c_loc(GRAD%DFEMBD)
    end if
    rc = msm_adjoint(gm_SIM, DLDF, DLDE, GRAD%POTR, GRAD%FPOTR, GRAD%POTB, GRAD%FPOTB, pf, pd)
    call msm_check(rc, "Adjoint_ForceTable_DEV", IERR)
  end subroutine Adjoint_ForceTable_DEV

  subroutine Clear_Globle_Variables_DEV()
    integer(c_int) :: rc
    if (c_associated(gm_SIM)) rc = msm_destroy(gm_SIM)
    gm_SIM = c_null_ptr
    gm_NBOX = 0
    gm_NPRT = 0
  end subroutine Clear_Globle_Variables_DEV

  ! ---------------------------------------------------------------------------------------------
  ! external forces (AddExtForce_ForceClass / Cumulate_ExtForce, MD_ForceClass_Register_GPU.F90:
  ! 584-680, 982-1180)
  ! ---------------------------------------------------------------------------------------------

  ! Register ForceArray / EpotArray under Tag (a tag already present is replaced); they run after
  ! the force pass of CalForce_ForceClass / the energy pass of CalEpot_ForceClass (For_One_Step
  ! included; the fused steps of Run_Steps_DEV do not call them)
  subroutine AddExtForce_ForceClass(ForceClass, Tag, ForceArray, EpotArray)
    type(MDForceClassGPU), intent(inout) :: ForceClass
    character(len=*), intent(in) :: Tag
    procedure(EXTFORCE_PROC), optional :: ForceArray, EpotArray
    integer :: k, slot
    slot = 0
    do k = 1, ForceClass%NEXT
      if (ForceClass%EXT(k)%Tag == Tag) slot = k
    end do
    if (slot == 0) then
      if (ForceClass%NEXT >= mp_MXEXTFORCE) then
        write(*, "(A)") "MDPSCU Error: too many external forces registered"
        stop 1
      end if
      ForceClass%NEXT = ForceClass%NEXT + 1
      slot = ForceClass%NEXT
    end if
    ForceClass%EXT(slot)%Tag = Tag
    ForceClass%EXT(slot)%pForceArray => null()
    ForceClass%EXT(slot)%pEpotArray => null()
    if (present(ForceArray)) ForceClass%EXT(slot)%pForceArray => ForceArray
    if (present(EpotArray)) ForceClass%EXT(slot)%pEpotArray => EpotArray
  end subroutine AddExtForce_ForceClass

  ! Clear_ExtForce: drop the entry of Tag, or every entry when Tag is absent
  subroutine Clear_ExtForce(ForceClass, Tag)
    type(MDForceClassGPU), intent(inout) :: ForceClass
    character(len=*), optional, intent(in) :: Tag
    integer :: k, m
    if (.not. present(Tag)) then
      ForceClass%NEXT = 0
      return
    end if
    m = 0
    do k = 1, ForceClass%NEXT
      if (ForceClass%EXT(k)%Tag == Tag) cycle
      m = m + 1
      if (m /= k) then
        ForceClass%EXT(m)%Tag = ForceClass%EXT(k)%Tag
        ForceClass%EXT(m)%pForceArray => ForceClass%EXT(k)%pForceArray
        ForceClass%EXT(m)%pEpotArray => ForceClass%EXT(k)%pEpotArray
      end if
    end do
    ForceClass%NEXT = m
  end subroutine Clear_ExtForce

  ! FEXT(NPRT,3) of all boxes, in the caller's atom order, added to the device forces
  subroutine AddForceFrom_Host_to_Devices(FEXT, IERR)
    real(KINDDF), intent(in), target, contiguous :: FEXT(:, :)
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc
    if (size(FEXT, 1) /= gm_NPRT .or. size(FEXT, 2) /= 3) then
      rc = MSM_ERR_ARG
    else
      rc = msm_add_forces(gm_SIM, c_loc(FEXT), 0_c_int)
    end if
    call msm_check(rc, "AddForceFrom_Host_to_Devices", IERR)
  end subroutine AddForceFrom_Host_to_Devices

  ! EEXT(NPRT) added to the device's per-atom energies
  subroutine AddEpotFrom_Host_to_Devices(EEXT, IERR)
    real(KINDDF), intent(in), target, contiguous :: EEXT(:)
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc
    if (size(EEXT) /= gm_NPRT) then
      rc = MSM_ERR_ARG
    else
      rc = msm_add_epot(gm_SIM, c_loc(EEXT), 0_c_int)
    end if
    call msm_check(rc, "AddEpotFrom_Host_to_Devices", IERR)
  end subroutine AddEpotFrom_Host_to_Devices

  ! DIS of all boxes onto the device (a restart: Initialize_Globle_Variables_DEV starts DIS at 0)
  subroutine CopyDISFrom_Host_to_Devices(SimBox, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    integer, optional, intent(out) :: IERR
    real(c_double), allocatable, target :: a(:)
    integer :: n, b, off, k, m
    integer(c_int) :: rc
    n = gm_NPRT
    allocate(a(3 * max(n, 1)))
    a = 0.0d0
    off = 0
    do b = 1, size(SimBox)
      m = SimBox(b)%NPRT
      if (allocated(SimBox(b)%DIS)) then
        do k = 1, 3
          a((k - 1) * n + off + 1:(k - 1) * n + off + m) = SimBox(b)%DIS(:, k)
        end do
      end if
      off = off + m
    end do
    if (off /= n) then
      rc = MSM_ERR_ARG
    else
      rc = msm_put_dis(gm_SIM, c_loc(a))
    end if
    call msm_check(rc, "CopyDISFrom_Host_to_Devices", IERR)
  end subroutine CopyDISFrom_Host_to_Devices

  ! Thermalizing_MC_DEV (MD_DiffScheme_GPU.F90:1451-1805): Maxwell velocities at TI (K) on the
  ! device, components with FIXVEL/FIXPOS kept, inactive atoms at rest, each box's mass-weighted
  ! mean velocity removed; the device generator is keyed by CtrlParam%SEED(1:2) and advances with
  ! every call (msm_thermalize)
  subroutine Thermalizing_MC_DEV(SimBox, CtrlParam, TI, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    real(KINDDF), intent(in) :: TI
    integer, optional, intent(out) :: IERR
    integer(c_long_long) :: seed
    seed = ior(ishft(int(CtrlParam%SEED(2), c_long_long), 32), &
               iand(int(CtrlParam%SEED(1), c_long_long), int(z'FFFFFFFF', c_long_long)))
    call msm_check(msm_thermalize(gm_SIM, TI, CtrlParam%KB, seed), "Thermalizing_MC_DEV", IERR)
  end subroutine Thermalizing_MC_DEV

  ! ---------------------------------------------------------------------------------------------
  ! binary restart records (Archive/Restore_Config_SimBoxArray, MD_SimBoxArray.F90:1964-2047):
  ! the stamp (Archive_RecordStamp, MD_TypeDef_RecordStamp.F90:195-226), then per box the records
  ! of Archive_Config_SimMDBox0 (MD_TypeDef_SimBox.F90:3268-3304) and an empty data pad
  ! (Archive_DataPad, MSMLIB/sor/Common/MSM_TypeDef_DataPad.F90:2925-2973), then the caller's
  ! own records (ExtArchive / ExtRestore, optional); Fortran unformatted sequential, unit chosen
  ! free from 10 up
  ! ---------------------------------------------------------------------------------------------

  integer function free_unit() result(u)
    logical :: op
    do u = 10, 999
      inquire(unit=u, opened=op)
      if (.not. op) return
    end do
    u = -1
  end function free_unit

  subroutine Archive_Config_SimBoxArray(SimBox, Stamp, fname, ExtArchive)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(MDRecordStamp), intent(in) :: Stamp
    character(len=*), intent(in) :: fname
    interface
      subroutine ExtArchive(hFile)
        integer, intent(in) :: hFile
      end subroutine ExtArchive
    end interface
    optional :: ExtArchive
    integer :: hf, nb, i, np, nxp, ng
    real(KINDDF) :: cm(mp_MXGROUP)
    integer :: ptype(mp_MXGROUP, mp_MXGROUP)
    real(KINDDF), allocatable :: z3(:, :), z1(:)
    character(len=64) :: tag
    hf = free_unit()
    nb = size(SimBox)
    if (Stamp%NBox > 0) nb = Stamp%NBox
    open(hf, file=fname, form="unformatted", status="unknown")
    write(hf) Stamp%AppType(1:len(Stamp%AppType))
    write(hf) Stamp%ITest, Stamp%NBox, Stamp%IBox, Stamp%ICfg, Stamp%IRec, Stamp%ITime, &
              Stamp%ISect, Stamp%Time, Stamp%ScalTime, Stamp%InstantTemp
    do i = 1, nb
      associate(B => SimBox(i))
        ng = B%NGROUP
        cm = 0.0d0
        ptype = 0
        if (allocated(B%CM)) cm(1:size(B%CM)) = B%CM
        if (allocated(B%PTYPE)) ptype(1:size(B%PTYPE, 1), 1:size(B%PTYPE, 2)) = B%PTYPE
        write(hf) B%RR, B%LATT, B%BOXSHAPE, B%BOXLOW, B%BOXLOW + B%ZL, B%ZL
        write(hf) B%NPRT, B%NGROUP
        write(hf) B%NA, B%IPA, B%SYMB, B%CZ, cm, B%PROP, ptype
        write(hf) B%POTTYPE, B%POTLIBNAME, B%POTSUBLIBNAME
        write(hf) B%TEMPERATURE, B%VOLUME, B%SPRESS0, B%SPRESS1, B%SPRESS, B%AVEPOT, B%HARMIL
        write(hf) B%VTENSOR, B%KTENSOR, B%PTENSOR
        np = B%NPRT
        nxp = 1
        allocate(z3(np, 3), z1(np))
        z3 = 0.0d0
        z1 = 0.0d0
        write(hf) np, nxp
        write(hf) B%ITYP(1:np)
        write(hf) B%XP(1:np, 1:3)
        write(hf) B%XP1(1:np, 1:3)
        if (allocated(B%FP)) then
          write(hf) B%FP(1:np, 1:3)
        else
          write(hf) z3
        end if
        if (allocated(B%EKIN)) then
          write(hf) B%EKIN(1:np)
        else
          write(hf) z1
        end if
        if (allocated(B%EPOT)) then
          write(hf) B%EPOT(1:np)
        else
          write(hf) z1
        end if
        write(hf) B%STATU(1:np)
        if (allocated(B%DIS)) then
          write(hf) B%DIS(1:np, 1:3)
        else
          write(hf) z3
        end if
        deallocate(z3, z1)
      end associate
      tag = "PADDATA"
      write(hf) tag
      write(hf) 0
    end do
    if (present(ExtArchive)) call ExtArchive(hf)
    close(hf)
  end subroutine Archive_Config_SimBoxArray

  subroutine Restore_Config_SimBoxArray(SimBox, Stamp, fname, ExtRestore)
    type(SimMDBox), intent(inout) :: SimBox(:)
    type(MDRecordStamp), intent(inout) :: Stamp
    character(len=*), intent(in) :: fname
    interface
      subroutine ExtRestore(hFile)
        integer, intent(in) :: hFile
      end subroutine ExtRestore
    end interface
    optional :: ExtRestore
    integer :: hf, nb, i, np, nxp, ng, npad, ios
    real(KINDDF) :: cm(mp_MXGROUP), boxup(3)
    integer :: ptype(mp_MXGROUP, mp_MXGROUP)
    real(KINDDF), allocatable :: skip3(:, :)
    character(len=64) :: tag
    hf = free_unit()
    nb = size(SimBox)
    if (Stamp%NBox > 0) nb = Stamp%NBox
    open(hf, file=fname, form="unformatted", status="old", iostat=ios)
    if (ios /= 0) then
      write(*, "(A)") "MDPSUC error: fail to restore the box: cannot open " // trim(fname)
      stop 1
    end if
    read(hf, iostat=ios) Stamp%AppType(1:len(Stamp%AppType))
    if (ios == 0) read(hf, iostat=ios) Stamp%ITest, Stamp%NBox, Stamp%IBox, Stamp%ICfg, &
                                       Stamp%IRec, Stamp%ITime, Stamp%ISect, Stamp%Time, &
                                       Stamp%ScalTime, Stamp%InstantTemp
    do i = 1, nb
      if (ios /= 0) exit
      associate(B => SimBox(i))
        read(hf, iostat=ios) B%RR, B%LATT, B%BOXSHAPE, B%BOXLOW, boxup, B%ZL
        if (ios == 0) read(hf, iostat=ios) B%NPRT, B%NGROUP
        if (ios == 0) read(hf, iostat=ios) B%NA, B%IPA, B%SYMB, B%CZ, cm, B%PROP, ptype
        if (ios == 0) read(hf, iostat=ios) B%POTTYPE, B%POTLIBNAME, B%POTSUBLIBNAME
        if (ios == 0) read(hf, iostat=ios) B%TEMPERATURE, B%VOLUME, B%SPRESS0, B%SPRESS1, &
                                           B%SPRESS, B%AVEPOT, B%HARMIL
        if (ios == 0) read(hf, iostat=ios) B%VTENSOR, B%KTENSOR, B%PTENSOR
        if (ios == 0) read(hf, iostat=ios) np, nxp
        if (ios /= 0) exit
        ng = max(B%NGROUP, 1)
        if (allocated(B%CM)) deallocate(B%CM)
        if (allocated(B%PTYPE)) deallocate(B%PTYPE)
        allocate(B%CM(ng), B%PTYPE(ng, ng))
        B%CM = cm(1:ng)
        B%PTYPE = ptype(1:ng, 1:ng)
        if (allocated(B%ITYP)) deallocate(B%ITYP)
        if (allocated(B%STATU)) deallocate(B%STATU)
        if (allocated(B%XP)) deallocate(B%XP)
        if (allocated(B%XP1)) deallocate(B%XP1)
        if (allocated(B%FP)) deallocate(B%FP)
        if (allocated(B%DIS)) deallocate(B%DIS)
        if (allocated(B%EKIN)) deallocate(B%EKIN)
        if (allocated(B%EPOT)) deallocate(B%EPOT)
        allocate(B%ITYP(np), B%STATU(np), B%XP(np, 3), B%XP1(np, 3), B%FP(np, 3), B%DIS(np, 3), &
                 B%EKIN(np), B%EPOT(np))
        read(hf, iostat=ios) B%ITYP(1:np)
        if (ios == 0) read(hf, iostat=ios) B%XP(1:np, 1:3)
        if (ios == 0) read(hf, iostat=ios) B%XP1(1:np, 1:3)
        if (ios == 0) read(hf, iostat=ios) B%FP(1:np, 1:3)
        if (ios == 0) read(hf, iostat=ios) B%EKIN(1:np)
        if (ios == 0) read(hf, iostat=ios) B%EPOT(1:np)
        if (ios == 0) read(hf, iostat=ios) B%STATU(1:np)
        if (ios == 0) read(hf, iostat=ios) B%DIS(1:np, 1:3)
        if (ios == 0 .and. nxp >= 5) then
          ! the higher-order Gear predictor arrays XP2..XP5 (not used by the Swope scheme)
          allocate(skip3(np, 3))
          read(hf, iostat=ios) skip3
          if (ios == 0) read(hf, iostat=ios) skip3
          if (ios == 0) read(hf, iostat=ios) skip3
          if (ios == 0) read(hf, iostat=ios) skip3
          deallocate(skip3)
        end if
      end associate
      ! Restore_DataPad: a data pad this layer does not keep is read past
      if (ios == 0) read(hf, iostat=ios) tag
      if (ios == 0) read(hf, iostat=ios) npad
      if (ios == 0 .and. npad /= 0) then
        write(*, "(A)") "MDPSUC error: data pads in the restart file are not supported"
        ios = 1
      end if
    end do
    if (ios /= 0) then
      write(*, "(A)") "MDPSUC error: fail to restore the box"
      close(hf)
      stop 1
    end if
    if (present(ExtRestore)) call ExtRestore(hf)
    close(hf)
  end subroutine Restore_Config_SimBoxArray

end module MD_ForceLib_HIP
