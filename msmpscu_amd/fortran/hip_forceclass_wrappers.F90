!  hip_forceclass_wrappers.F90 -- the glue a reference build adds to bind its force class to the
!  HIP path: wrappers with the signatures of the reference's force-class procedure pointers
!  (MDLIB/sor/CommonGPU/MD_ForceClass_Register_GPU.F90:29-226) that copy the fields the path reads
!  from the reference's SimMDBox / SimMDCtrl / MDForceTable into this layer's types of the same
!  names (renamed on `use`) and call the layer, and HIP_Register_ForceClass, the body of the one
!  extra case a maintainer adds to the reference's Register_ForceClass (:363-442).  The reference's
!  arrays are CGS, so the layer's Boltzmann constant is set to the reference's CP_KB.
module HIP_ForceClass_Wrappers
  use MD_CONSTANTS, only: KINDDF
  use MD_TYPEDEF_SimMDBox, only: SimMDBox
  use MD_TYPEDEF_SimMDCtrl, only: SimMDCtrl
  use MD_TYPEDEF_ForceTable, only: MDForceTable
  use MD_ForceClass_Register_GPU, only: MDForceClassGPU
  use MD_ForceLib_HIP, only: HBox => SimMDBox, HCtrl => SimMDCtrl, HClass => MDForceClassGPU, &
                             CP_KB_CGS, Register_ForceClass, Initialize_Globle_Variables_DEV, &
                             Init_Forcetable_Dev, Cal_NeighBoreList_DEV, CalForce_ForceClass, &
                             CalEpot_ForceClass, CalEDen_ForceClass, CalPTensor_ForceClass, &
                             Cal_AtomicStress, CopyOut_SimBox_DEV, Clear_Globle_Variables_DEV
  implicit none
  private
  public :: HIP_Register_ForceClass
  type(HBox), save :: hb(1)
  type(HCtrl), save :: hc
  type(HClass), save :: hf

contains

  ! the case "HIP_FS_TYPE" / "HIP_EAM_TYPE" of Register_ForceClass
  subroutine HIP_Register_ForceClass(ForceClass)
    type(MDForceClassGPU), intent(inout) :: ForceClass
    ForceClass%pIniForcetable => HIP_IniForcetable
    ForceClass%pClrForcetable => HIP_ClrForcetable
    ForceClass%pCalForce => HIP_CalForce
    ForceClass%pCalEpot0 => HIP_CalEpot
    ForceClass%pCalEpot => HIP_CalEpot
    ForceClass%pCalEDen => HIP_CalEDen
    ForceClass%pCalPTensor => HIP_CalPTensor
    ForceClass%pCalAVStress => HIP_CalAVStress
  end subroutine HIP_Register_ForceClass

  subroutine to_hip(SimBox, CtrlParam)
    type(SimMDBox), intent(in) :: SimBox
    type(SimMDCtrl), intent(in) :: CtrlParam
    integer :: ng
    ng = SimBox%NGROUP
    hb(1)%NPRT = SimBox%NPRT
    hb(1)%NGROUP = ng
    hb(1)%ZL = SimBox%ZL
    hb(1)%BOXLOW = SimBox%BOXLOW
    hb(1)%RR = SimBox%RR
    hb(1)%XP = SimBox%XP
    hb(1)%XP1 = SimBox%XP1
    hb(1)%ITYP = SimBox%ITYP
    hb(1)%STATU = SimBox%STATU
    hb(1)%CM = SimBox%CM(1:ng)
    hc%RU = CtrlParam%RU(1:ng, 1:ng)
    hc%NB_RM = CtrlParam%NB_RM(1:ng, 1:ng)
    hc%NB_UPTAB = CtrlParam%NB_UPTAB
    hc%NB_MXNBS = CtrlParam%NB_MXNBS
    hc%NUMFTABR = CtrlParam%NUMFTABR
    hc%H = CtrlParam%H
    hc%IFPD = CtrlParam%IFPD
    hc%KB = CP_KB_CGS
  end subroutine to_hip

  subroutine HIP_IniForcetable(SimBox, CtrlParam, FTable, RelaseTable, MULTIBOX)
    type(SimMDBox), intent(inout) :: SimBox
    type(SimMDCtrl), intent(in) :: CtrlParam
    type(MDForceTable), intent(in) :: FTable
    integer, optional :: RelaseTable, MULTIBOX
    call to_hip(SimBox, CtrlParam)
    call Register_ForceClass(trim(FTable%PotType), hf)
    ! the reference's tables, unchanged layout (NKIND, NTAB) and (NKIND1, NEMBD)
    hf%ForceTable%NKIND = FTable%NKIND
    hf%ForceTable%NTAB = FTable%NTAB
    hf%ForceTable%Rmax = FTable%Rmax
    hf%ForceTable%CSI = FTable%CSI
    hf%ForceTable%KPAIR = FTable%KPAIR
    hf%ForceTable%POTR = FTable%POTR
    hf%ForceTable%FPOTR = FTable%FPOTR
    hf%ForceTable%POTB = FTable%POTB
    hf%ForceTable%FPOTB = FTable%FPOTB
    if (trim(FTable%PotType) == "EAM_TYPE") then
      hf%ForceTable%NKIND1 = FTable%NKIND1
      hf%ForceTable%NEMBD = FTable%NEMBD
      hf%ForceTable%RHOD = FTable%RHOD
      hf%ForceTable%KEMBD = FTable%KEMBD
      hf%ForceTable%FEMBD = FTable%FEMBD
      hf%ForceTable%DFEMBD = FTable%DFEMBD
    end if
    call Initialize_Globle_Variables_DEV(hb, hc)
    call Init_Forcetable_Dev(hb, hc, hf)
    call Cal_NeighBoreList_DEV(hb, hc)
  end subroutine HIP_IniForcetable

  subroutine HIP_CalForce(SimBox, CtrlParam)
    type(SimMDBox), intent(inout) :: SimBox
    type(SimMDCtrl), intent(in) :: CtrlParam
    call CalForce_ForceClass(hb, hc, hf)
    call CopyOut_SimBox_DEV(hb)
    SimBox%FP = hb(1)%FP
  end subroutine HIP_CalForce

  subroutine HIP_CalEpot(SimBox, CtrlParam)
    type(SimMDBox), intent(inout) :: SimBox
    type(SimMDCtrl), intent(in) :: CtrlParam
    call CalEpot_ForceClass(hb, hc, hf)
    call CopyOut_SimBox_DEV(hb, WITHEPOT=.true.)
    SimBox%EPOT = hb(1)%EPOT
  end subroutine HIP_CalEpot

  ! F'(rho) stays on the device, as DEN does in the reference
  subroutine HIP_CalEDen(SimBox, CtrlParam)
    type(SimMDBox), intent(inout) :: SimBox
    type(SimMDCtrl), intent(in) :: CtrlParam
    call CalEDen_ForceClass(hb, hc, hf)
  end subroutine HIP_CalEDen

  subroutine HIP_CalPTensor(SimBox, CtrlParam)
    type(SimMDBox), intent(inout) :: SimBox
    type(SimMDCtrl), intent(in) :: CtrlParam
    real(KINDDF) :: vt(3, 3)
    call CalPTensor_ForceClass(hb, hc, hf, vt)
    SimBox%VTENSOR = vt
  end subroutine HIP_CalPTensor

  ! AVP(NPRT, 9) in the caller's atom order (a host array here; the reference's is on the device)
  subroutine HIP_CalAVStress(IDEV, dAVP)
    integer, intent(in) :: IDEV
    real(KINDDF), dimension(:, :), intent(out) :: dAVP
    call Cal_AtomicStress(hb, hc, dAVP)
  end subroutine HIP_CalAVStress

  subroutine HIP_ClrForcetable()
    call Clear_Globle_Variables_DEV()
  end subroutine HIP_ClrForcetable

end module HIP_ForceClass_Wrappers
