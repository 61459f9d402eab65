!  mdlib_ref_types.F90 -- the part of the reference's MDLIB type modules that the force-class glue
!  (hip_forceclass_wrappers.F90) reads, under the reference's module names, type names, field names,
!  kinds and shapes, so that the glue compiles here exactly as it would inside a reference build.
!  It is not the reference: only the fields the path reads are declared, with the reference's
!  defaults, and values are in the reference's CGS units (cm, g, s, erg).
!
!    MD_CONSTANTS           KINDDF, mp_MXGROUP          (MDLIB/sor/Common/MD_Const.F90:78)
!    MD_TYPEDEF_SimMDBox    type SimMDBox               (MDLIB/sor/Common/MD_TypeDef_SimBox.F90:43-132)
!    MD_TYPEDEF_SimMDCtrl   type SimMDCtrl              (MDLIB/sor/Common/MD_TypeDef_SimCtrlParam.F90:56-142)
!    MD_TYPEDEF_ForceTable  type MDForceTable           (MDLIB/sor/Common/MD_TypeDef_ForceTable.F90:117-155)
!    MD_ForceClass_Register_GPU  the abstract interfaces and the procedure-pointer type
!                           MDForceClassGPU             (MDLIB/sor/CommonGPU/MD_ForceClass_Register_GPU.F90:29-259)
!  The one change: CALATOMICSTRESS's dAVP is a host array here (the reference declares it with
!  CUDA Fortran's `device` attribute, which no compiler in this image accepts).

module MD_CONSTANTS
  implicit none
  integer, parameter :: KINDDF = kind(1.0d0)
  integer, parameter :: mp_MXGROUP = 16
end module MD_CONSTANTS

module MD_TYPEDEF_SimMDBox
  use MD_CONSTANTS
  implicit none
  type :: SimMDBox
    real(KINDDF) :: RR = -1.d0
    real(KINDDF) :: ZL(3) = 0.d0
    real(KINDDF) :: BOXLOW(3) = 1.d64
    real(KINDDF) :: BOXUP(3) = -1.d64
    integer :: NPRT = 0
    integer :: NGROUP = 0
    integer, dimension(mp_MXGROUP) :: NA = 0
    real(KINDDF), dimension(mp_MXGROUP) :: CM = 0.d0
    integer, dimension(:), allocatable :: ITYP
    real(KINDDF), dimension(:, :), allocatable :: XP
    real(KINDDF), dimension(:, :), allocatable :: XP1
    real(KINDDF), dimension(:, :), allocatable :: DIS
    integer, dimension(:), allocatable :: STATU
    real(KINDDF), dimension(:, :), allocatable :: FP
    real(KINDDF), dimension(:), allocatable :: EPOT
    real(KINDDF) :: VTENSOR(3, 3) = 0.d0
  end type SimMDBox
end module MD_TYPEDEF_SimMDBox

module MD_TYPEDEF_SimMDCtrl
  use MD_CONSTANTS
  implicit none
  type :: SimMDCtrl
    real(KINDDF) :: H = 0.d0
    integer :: IFPD(3) = 1
    integer :: NB_MXNBS = 800
    real(KINDDF) :: NB_RM(mp_MXGROUP, mp_MXGROUP) = 0.d0
    integer :: NB_UPTAB = 10
    real(KINDDF) :: RU(mp_MXGROUP, mp_MXGROUP) = 0.d0
    integer :: NUMFTABR = 10000
  end type SimMDCtrl
end module MD_TYPEDEF_SimMDCtrl

module MD_TYPEDEF_ForceTable
  use MD_CONSTANTS
  implicit none
  type :: MDForceTable
    character*12 :: PotType = "FS_TYPE"
    integer :: NKIND = 0
    integer :: NTAB = 10000
    real(KINDDF) :: Rmax = -1
    real(KINDDF) :: CSI
    integer, dimension(:, :), allocatable :: KPAIR
    real(KINDDF), dimension(:, :), allocatable :: POTR
    real(KINDDF), dimension(:, :), allocatable :: FPOTR
    real(KINDDF), dimension(:, :), allocatable :: POTB
    real(KINDDF), dimension(:, :), allocatable :: FPOTB
    integer :: NEMBD = 10000
    integer :: NKIND1 = 0
    integer, dimension(:), allocatable :: KEMBD
    real(KINDDF), dimension(:, :), allocatable :: FEMBD
    real(KINDDF), dimension(:, :), allocatable :: DFEMBD
    real(KINDDF) :: RHOMX = -1
    real(KINDDF) :: RHOD = 1
  end type MDForceTable
end module MD_TYPEDEF_ForceTable

module MD_ForceClass_Register_GPU
  use MD_CONSTANTS
  use MD_TYPEDEF_SimMDBox
  use MD_TYPEDEF_SimMDCtrl
  use MD_TYPEDEF_ForceTable
  implicit none
  abstract interface
    subroutine INITIALIZE_FORCETABLE(SimBox, CtrlParam, FTable, RelaseTable, MULTIBOX)
      import :: SimMDBox, SimMDCtrl, MDForceTable
      type(SimMDBox), intent(inout) :: SimBox
      type(SimMDCtrl), intent(in) :: CtrlParam
      type(MDForceTable), intent(in) :: FTable
      integer, optional :: RelaseTable, MULTIBOX
    end subroutine INITIALIZE_FORCETABLE
    subroutine CLEAR_FORCETABLE()
    end subroutine CLEAR_FORCETABLE
    subroutine CALFORCE(SimBox, CtrlParam)
      import :: SimMDBox, SimMDCtrl
      type(SimMDBox), intent(inout) :: SimBox
      type(SimMDCtrl), intent(in) :: CtrlParam
    end subroutine CALFORCE
    subroutine CALEPOT(SimBox, CtrlParam)
      import :: SimMDBox, SimMDCtrl
      type(SimMDBox), intent(inout) :: SimBox
      type(SimMDCtrl), intent(in) :: CtrlParam
    end subroutine CALEPOT
    subroutine CALEDEN(SimBox, CtrlParam)
      import :: SimMDBox, SimMDCtrl
      type(SimMDBox), intent(inout) :: SimBox
      type(SimMDCtrl), intent(in) :: CtrlParam
    end subroutine CALEDEN
    subroutine CALPTENSOR(SimBox, CtrlParam)
      import :: SimMDBox, SimMDCtrl
      type(SimMDBox), intent(inout) :: SimBox
      type(SimMDCtrl), intent(in) :: CtrlParam
    end subroutine CALPTENSOR
    subroutine CALATOMICSTRESS(IDEV, dAVP)
      import :: KINDDF
      integer, intent(in) :: IDEV
      real(KINDDF), dimension(:, :), intent(out) :: dAVP
    end subroutine CALATOMICSTRESS
  end interface
  type :: MDForceClassGPU
    procedure(INITIALIZE_FORCETABLE), pointer, nopass :: pIniForcetable => null()
    procedure(CLEAR_FORCETABLE), pointer, nopass :: pClrForcetable => null()
    procedure(CALFORCE), pointer, nopass :: pCalForce => null()
    procedure(CALEPOT), pointer, nopass :: pCalEpot0 => null()
    procedure(CALEPOT), pointer, nopass :: pCalEpot => null()
    procedure(CALEDEN), pointer, nopass :: pCalEDen => null()
    procedure(CALPTENSOR), pointer, nopass :: pCalPTensor => null()
    procedure(CALATOMICSTRESS), pointer, nopass :: pCalAVStress => null()
    type(MDForceTable) :: ForceTable
  end type MDForceClassGPU
end module MD_ForceClass_Register_GPU
