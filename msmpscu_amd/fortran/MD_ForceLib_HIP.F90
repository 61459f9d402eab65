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

  ! --- SimMDBox (subset of MD_TypeDef_SimBox.F90:43-132 used by the force path)
  type, public :: SimMDBox
    integer :: NPRT = 0, NGROUP = 1
    real(KINDDF) :: ZL(3) = 0.0d0, BOXLOW(3) = 0.0d0
    real(KINDDF) :: RR = 1.0d0   ! lattice unit, in this layer's length unit (Angstrom)
    real(KINDDF), allocatable :: XP(:, :), XP1(:, :), FP(:, :), DIS(:, :), EPOT(:), CM(:)
    integer, allocatable :: ITYP(:), STATU(:), PTYPE(:, :)
  end type SimMDBox

  ! --- SimMDCtrl (fields of MD_TypeDef_SimCtrlParam.F90:56-142 used by the force path)
  type, public :: SimMDCtrl
    real(KINDDF), allocatable :: RU(:, :), NB_RM(:, :)
    integer :: NB_UPTAB = 10, NB_MXNBS = 800, NUMFTABR = 10000, NUMFTABE = 10000
    real(KINDDF) :: RHOSCAL = 20.0d0, H = 1.0d0
    ! quench controls (MD_TypeDef_SimCtrlParam.F90:213-217): moves in lattice units RR, energy in eV
    real(KINDDF) :: STEEPEST_Alpha = 0.1d0, STEEPEST_MiStep = 1.0d-5, STEEPEST_MxStep = 0.1d0
    real(KINDDF) :: STEEPEST_MiDelE = 1.0d-3
    integer :: IFPD(3) = 1
  end type SimMDCtrl

  ! --- MDForceClassGPU (MD_ForceClass_Register_GPU.F90:248-259): the class name selects the
  !     FS or EAM device passes; the table travels with it
  type, public :: MDForceClassGPU
    character(len=16) :: CLASSNAME = ""
    type(MDForceTable) :: ForceTable
  end type MDForceClassGPU

  ! device context of this process (the reference keeps its device state in module variables too)
  type(c_ptr), save :: gm_SIM = c_null_ptr
  integer, save :: gm_NBOX = 0, gm_NPRT = 0
  integer, allocatable, save :: gm_NAT(:)

  public :: Register_ForceTableProc, Create_Interaction_ForceTable, Register_ForceClass
  public :: Initialize_Globle_Variables_DEV, Init_Forcetable_Dev, Cal_NeighBoreList_DEV
  public :: CalForce_ForceClass, CalEpot_ForceClass, CalPTensor_ForceClass, Cal_AtomicStress
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

  ! Register_ForceTableProc (type-ID form, MD_TypeDef_ForceTable.F90:626-730)
  subroutine Register_ForceTableProc(IFORCE, FTable, NNFORCE, NEFORCE, EMBDF, NOTE)
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
  end subroutine Register_ForceTableProc

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
    type(SimMDBox), intent(in) :: SimBox
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
  subroutine CalForce_ForceClass(SimBox, CtrlParam, ForceClass, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    integer, optional, intent(out) :: IERR
    call msm_check(msm_calc_force(gm_SIM), "CalForce_ForceClass", IERR)
  end subroutine CalForce_ForceClass

  ! CalEpot_ForceClass (:778-822)
  subroutine CalEpot_ForceClass(SimBox, CtrlParam, ForceClass, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    integer, optional, intent(out) :: IERR
    call msm_check(msm_calc_epot(gm_SIM), "CalEpot_ForceClass", IERR)
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
    rc = msm_calc_ekin(gm_SIM, CP_KB_EV, ek, tb)
    call msm_check(rc, "Cal_GlobalT_DEV", IERR)
    if (rc /= MSM_OK) return
    if (count(ek(1:gm_NPRT) >= 0.0d0) > 0) &
      CURT = 2.0d0 * sum(ek(1:gm_NPRT), mask=(ek(1:gm_NPRT) >= 0.0d0)) / &
             dble(count(ek(1:gm_NPRT) >= 0.0d0)) / (3.0d0 * CP_KB_EV)
  end subroutine Cal_GlobalT_DEV

  ! VelScaling_DEV (MD_DiffScheme_GPU.F90:1390-1446): scale the velocities of each box to DT (K)
  subroutine VelScaling_DEV(SimBox, CtrlParam, DT, IERR)
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    real(KINDDF), intent(in) :: DT
    integer, optional, intent(out) :: IERR
    call msm_check(msm_scale_velocities(gm_SIM, DT, CP_KB_EV), "VelScaling_DEV", IERR)
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

  ! Predictor_DEV / Correction_DEV (MD_DiffScheme_GPU.F90:604-682 / 821-847)
  subroutine Predictor_DEV(ITIME, SimBox, CtrlParam, IERR)
    integer, intent(in) :: ITIME
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    integer, optional, intent(out) :: IERR
    call msm_check(msm_predictor(gm_SIM, CtrlParam%H), "Predictor_DEV", IERR)
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
  subroutine For_One_Step(ITIME, SimBox, CtrlParam, ForceClass, IERR)
    integer, intent(in) :: ITIME
    type(SimMDBox), intent(in) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceClassGPU), intent(in) :: ForceClass
    integer, optional, intent(out) :: IERR
    integer(c_int) :: rc
    rc = msm_predictor(gm_SIM, CtrlParam%H)
    if (rc == MSM_OK .and. CtrlParam%NB_UPTAB > 0) then
      if (mod(ITIME, CtrlParam%NB_UPTAB) == 0) rc = msm_build_neighbors(gm_SIM)
    end if
    if (rc == MSM_OK) rc = msm_calc_force(gm_SIM)
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

end module MD_ForceLib_HIP
