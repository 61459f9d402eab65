!  ref_glue_driver.F90 -- a caller written against the reference's own types (mdlib_ref_types.F90:
!  SimMDBox, SimMDCtrl, MDForceTable and the procedure-pointer MDForceClassGPU, in CGS), bound to
!  the HIP path through hip_forceclass_wrappers.F90 and driven only through the force class's
!  pointers, as the reference's application code drives its own: pIniForcetable, pCalEDen,
!  pCalForce, pCalEpot, pCalPTensor, pCalAVStress, pClrForcetable.  Its W-W table is made the way a
!  reference program makes it (pair and density routines tabulated on the reference's grid) and
!  then handed over in the reference's table type.
!
!  usage: ref_glue_driver OUTFILE [NC]
!  OUTFILE (CGS): "N E0"; N lines "x y z fx fy fz den avp_trace"; "# vtensor" and nine values
!  (column-major) of SimBox%VTENSOR.
program ref_glue_driver
  use iso_c_binding
  use MD_CONSTANTS, only: KINDDF
  use MD_TYPEDEF_SimMDBox, only: SimMDBox
  use MD_TYPEDEF_SimMDCtrl, only: SimMDCtrl
  use MD_TYPEDEF_ForceTable, only: MDForceTable
  use MD_ForceClass_Register_GPU, only: MDForceClassGPU
  use HIP_ForceClass_Wrappers, only: HIP_Register_ForceClass
  use MSM_HIP_C, only: msm_drand32
  use MD_ForceLib_HIP, only: TBox => SimMDBox, TCtrl => SimMDCtrl, TClass => MDForceClassGPU, &
                             Register_ForceClass, Register_ForceTableProc, &
                             Create_Interaction_ForceTable, CopyDENFrom_Devices_to_Host
  use W_FS_CGS, only: W_NN_CGS, W_NE_CGS
  implicit none
  real(KINDDF), parameter :: AW = 3.165d-8, SIGMA = 0.03d-8
  real(KINDDF), parameter :: AMU_G = 1.66054d-24
  real(KINDDF), parameter :: TWOPI = 6.283185307179586d0
  type(SimMDBox) :: box
  type(SimMDCtrl) :: ctrl
  type(MDForceClassGPU) :: gm_ForceClass
  type(TBox) :: tb(1)
  type(TCtrl) :: tc
  type(TClass) :: tf
  character(len=512) :: outfile, arg
  integer :: nc, n, i, j, k, l, m, u
  integer(c_int) :: seed(2)
  real(KINDDF) :: z1, z2, x
  real(KINDDF), allocatable :: den(:), avp(:, :)

  if (command_argument_count() < 1) then
    write(*, "(A)") "usage: ref_glue_driver OUTFILE [NC]"
    stop 2
  end if
  call get_command_argument(1, outfile)
  nc = 8
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) nc
  end if

  ! --- the box in the reference's type, CGS
  n = 2 * nc * nc * nc
  box%NPRT = n
  box%NGROUP = 1
  box%NA(1) = n
  box%ZL = nc * AW
  box%RR = AW
  box%BOXLOW = -0.5d0 * box%ZL
  box%BOXUP = box%BOXLOW + box%ZL
  box%CM(1) = 183.84d0 * AMU_G
  allocate(box%XP(n, 3), box%XP1(n, 3), box%ITYP(n), box%STATU(n), box%FP(n, 3), box%EPOT(n))
  box%ITYP = 1
  box%STATU = 1
  m = 0
  do i = 0, nc - 1
    do j = 0, nc - 1
      do k = 0, nc - 1
        do l = 0, 1
          m = m + 1
          box%XP(m, 1) = (i + 0.5d0 * l) * AW + box%BOXLOW(1)
          box%XP(m, 2) = (j + 0.5d0 * l) * AW + box%BOXLOW(2)
          box%XP(m, 3) = (k + 0.5d0 * l) * AW + box%BOXLOW(3)
        end do
      end do
    end do
  end do
  seed = [int(24680, c_int), int(1357, c_int)]
  do k = 1, 3
    do m = 1, n
      z1 = msm_drand32(seed)
      z2 = msm_drand32(seed)
      x = box%XP(m, k) + SIGMA * sqrt(-2.0d0 * log(max(z1, 1.0d-300))) * cos(TWOPI * z2)
      x = box%BOXLOW(k) + modulo(x - box%BOXLOW(k), box%ZL(k))
      if (x >= box%BOXLOW(k) + box%ZL(k)) x = x - box%ZL(k)
      box%XP(m, k) = x
    end do
  end do
  box%XP1 = 0.0d0

  ! --- the control block in the reference's type
  ctrl%RU(1, 1) = 4.4002d-8
  ctrl%NB_RM(1, 1) = 4.7002d-8
  ctrl%NB_UPTAB = 10
  ctrl%NB_MXNBS = 64
  ctrl%NUMFTABR = 5000
  ctrl%H = 1.0d-15
  ctrl%IFPD = 1

  ! --- the W-W table on the reference's grid, handed over in the reference's table type
  tb(1)%NPRT = n
  tb(1)%NGROUP = 1
  tb(1)%ZL = box%ZL
  tb(1)%BOXLOW = box%BOXLOW
  tb(1)%SYMB(1) = "W"
  allocate(tb(1)%PTYPE(1, 1))
  tb(1)%PTYPE = 1
  allocate(tc%RU(1, 1), tc%NB_RM(1, 1))
  tc%RU = ctrl%RU(1, 1)
  tc%NB_RM = ctrl%NB_RM(1, 1)
  tc%NUMFTABR = ctrl%NUMFTABR
  call Register_ForceClass("FS_TYPE", tf)
  call Register_ForceTableProc(1, tf%ForceTable, W_NN_CGS, W_NE_CGS, NOTE="W-W FS, CGS")
  call Create_Interaction_ForceTable(tb(1), tc, tf%ForceTable)
  gm_ForceClass%ForceTable%PotType = "FS_TYPE"
  gm_ForceClass%ForceTable%NKIND = tf%ForceTable%NKIND
  gm_ForceClass%ForceTable%NTAB = tf%ForceTable%NTAB
  gm_ForceClass%ForceTable%Rmax = tf%ForceTable%Rmax
  gm_ForceClass%ForceTable%CSI = tf%ForceTable%CSI
  gm_ForceClass%ForceTable%KPAIR = tf%ForceTable%KPAIR
  gm_ForceClass%ForceTable%POTR = tf%ForceTable%POTR
  gm_ForceClass%ForceTable%FPOTR = tf%ForceTable%FPOTR
  gm_ForceClass%ForceTable%POTB = tf%ForceTable%POTB
  gm_ForceClass%ForceTable%FPOTB = tf%ForceTable%FPOTB

  ! --- bind and drive through the pointers only
  call HIP_Register_ForceClass(gm_ForceClass)
  call gm_ForceClass%pIniForcetable(box, ctrl, gm_ForceClass%ForceTable)
  call gm_ForceClass%pCalEDen(box, ctrl)
  allocate(den(n), avp(n, 9))
  call CopyDENFrom_Devices_to_Host(den)
  call gm_ForceClass%pCalForce(box, ctrl)
  call gm_ForceClass%pCalEpot(box, ctrl)
  call gm_ForceClass%pCalPTensor(box, ctrl)
  call gm_ForceClass%pCalAVStress(0, avp)

  open(newunit=u, file=trim(outfile), status="replace", action="write")
  write(u, "(I0,1X,ES25.17E3)") n, sum(box%EPOT)
  do m = 1, n
    write(u, "(8(ES25.17E3,1X))") box%XP(m, :), box%FP(m, :), den(m), &
        avp(m, 1) + avp(m, 5) + avp(m, 9)
  end do
  write(u, "(A)") "# vtensor"
  write(u, "(9(ES25.17E3,1X))") box%VTENSOR
  close(u)
  call gm_ForceClass%pClrForcetable()
end program ref_glue_driver
