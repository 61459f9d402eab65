!  cgs_driver.F90 -- a Fortran caller of the drop-in layer that works the way the reference's
!  MDLIB programs do: every quantity in CGS (lengths in cm, energies in erg, masses in g, time in
!  s, temperatures in K with CtrlParam%KB = CP_KB_CGS), the W-W potential registered by element
!  symbol ("W<-W") with routines in CGS, and the call surface around the force class:
!  CalEDen_ForceClass (density pass alone), CalForce / CalEpot, Thermalizing_MC_DEV and
!  Cal_GlobalT_DEV, an external force through AddExtForce_ForceClass, Predictor_DEV with its
!  damping window and variable-step search, and a binary restart (Archive / Restore of the
!  configuration records) resumed on a fresh device context.
!
!  usage: cgs_driver OUTFILE [NC]
!  OUTFILE (all values in CGS):
!    "N E0"; N lines "x y z fx fy fz den" at the start (den: F'(rho) of CalEDen_ForceClass)
!    "# temperature Tdev Thost" after Thermalizing_MC_DEV(600 K): Cal_GlobalT_DEV and 2 EK / 3 N kB
!      of the velocities copied out
!    "# extforce F0 maxdiff": constant force F0 along x added by a registered ForceArray routine;
!      max |F(with) - F(without) - F0 e_x| over atoms and components
!    "# predictor-in" and N lines "vx vy vz fx fy fz" before Predictor_DEV(ITIME = 1, damping
!      window 1..1, IHDUP = -1, HMX = 2 fs, DMX = 0.002 A), "# predictor H" the step it chose, N
!      lines "x y z vx vy vz" after it
!    "# restart dx dv ddis df dstatu": largest differences between 30 uninterrupted steps and
!      20 steps, archive, restore on a new context, 10 more steps (0 when the resume is exact)

module Push_Force
  use MD_ForceLib_HIP
  implicit none
  real(KINDDF), save :: F0 = 0.0d0
contains
  ! an external ForceArray routine: a constant force F0 along x on every atom
  subroutine Push_ForceArray(SimBox, CtrlParam)
    type(SimMDBox), intent(inout) :: SimBox(:)
    type(SimMDCtrl), intent(in) :: CtrlParam
    real(KINDDF), allocatable :: f(:, :)
    integer :: n
    n = sum(SimBox(:)%NPRT)
    allocate(f(n, 3))
    f = 0.0d0
    f(:, 1) = F0
    call AddForceFrom_Host_to_Devices(f)
  end subroutine Push_ForceArray
end module Push_Force

program cgs_driver
  use iso_c_binding
  use MSM_HIP_C, only: msm_drand32
  use MD_ForceLib_HIP
  use W_FS_CGS
  use Push_Force
  implicit none
  real(KINDDF), parameter :: AW = 3.165d-8, SIGMA = 0.03d-8
  real(KINDDF), parameter :: AMU_G = 1.66054d-24
  real(KINDDF), parameter :: TWOPI = 6.283185307179586d0
  type(SimMDBox) :: box(1), box2(1)
  type(SimMDBox), allocatable :: ref(:)
  type(SimMDCtrl) :: ctrl
  type(MDForceClassGPU) :: fc
  type(MDRecordStamp) :: stamp, stamp2
  character(len=512) :: outfile, arg, rstfile
  integer :: nc, n, i, j, k, l, m, u
  integer(c_int) :: seed(2)
  real(KINDDF) :: e0, z1, z2, x, tdev, thost, ek, maxd
  real(KINDDF), allocatable :: den(:), f_plain(:, :)
  real(KINDDF) :: dres(5)

  if (command_argument_count() < 1) then
    write(*, "(A)") "usage: cgs_driver OUTFILE [NC]"
    stop 2
  end if
  call get_command_argument(1, outfile)
  nc = 8
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) nc
  end if
  rstfile = trim(outfile) // ".rst"

  ! --- box in cm, the interaction found by element symbol (PTYPE = 0)
  n = 2 * nc * nc * nc
  box(1)%NPRT = n
  box(1)%NGROUP = 1
  box(1)%ZL = nc * AW
  box(1)%RR = AW
  box(1)%BOXLOW = -0.5d0 * box(1)%ZL
  box(1)%SYMB(1) = "W"
  allocate(box(1)%XP(n, 3), box(1)%XP1(n, 3), box(1)%ITYP(n), box(1)%STATU(n), box(1)%CM(1))
  allocate(box(1)%PTYPE(1, 1))
  box(1)%ITYP = 1
  box(1)%STATU = 1
  box(1)%CM(1) = 183.84d0 * AMU_G
  box(1)%PTYPE(1, 1) = 0
  m = 0
  do i = 0, nc - 1
    do j = 0, nc - 1
      do k = 0, nc - 1
        do l = 0, 1
          m = m + 1
          box(1)%XP(m, 1) = (i + 0.5d0 * l) * AW + box(1)%BOXLOW(1)
          box(1)%XP(m, 2) = (j + 0.5d0 * l) * AW + box(1)%BOXLOW(2)
          box(1)%XP(m, 3) = (k + 0.5d0 * l) * AW + box(1)%BOXLOW(3)
        end do
      end do
    end do
  end do
  seed = [int(31337, c_int), int(4711, c_int)]
  do k = 1, 3
    do m = 1, n
      z1 = msm_drand32(seed)
      z2 = msm_drand32(seed)
      x = box(1)%XP(m, k) + SIGMA * sqrt(-2.0d0 * log(max(z1, 1.0d-300))) * cos(TWOPI * z2)
      x = box(1)%BOXLOW(k) + modulo(x - box(1)%BOXLOW(k), box(1)%ZL(k))
      if (x >= box(1)%BOXLOW(k) + box(1)%ZL(k)) x = x - box(1)%ZL(k)
      box(1)%XP(m, k) = x
    end do
  end do
  box(1)%XP1 = 0.0d0

  ! --- control block in CGS
  allocate(ctrl%RU(1, 1), ctrl%NB_RM(1, 1))
  ctrl%RU = 4.4002d-8
  ctrl%NB_RM = 4.7002d-8
  ctrl%NB_UPTAB = 10
  ctrl%NB_MXNBS = 64
  ctrl%NUMFTABR = 5000
  ctrl%NUMFTABE = 5000
  ctrl%H = 1.0d-15
  ctrl%IFPD = 1
  ctrl%KB = CP_KB_CGS
  ctrl%SEED = [97531, 8642, 0, 0]
  call Register_ForceClass("FS_TYPE", fc)
  call Register_ForceTableProc("W", "W", fc%ForceTable, W_NN_CGS, W_NE_CGS, NOTE="W-W FS, CGS")
  call Create_Interaction_ForceTable(box(1), ctrl, fc%ForceTable)
  if (box(1)%PTYPE(1, 1) /= SymbID2TypeID_ForceTableProc("W", "W", fc%ForceTable)) stop 4

  ! --- density pass alone, then forces and energy
  call Initialize_Globle_Variables_DEV(box, ctrl)
  call Init_Forcetable_Dev(box, ctrl, fc)
  call Cal_NeighBoreList_DEV(box, ctrl)
  allocate(den(n))
  call CalEDen_ForceClass(box, ctrl, fc)
  call CopyDENFrom_Devices_to_Host(den)
  call CalForce_ForceClass(box, ctrl, fc)
  call CalEpot_ForceClass(box, ctrl, fc)
  call CopyOut_SimBox_DEV(box, WITHEPOT=.true.)
  e0 = sum(box(1)%EPOT)
  open(newunit=u, file=trim(outfile), status="replace", action="write")
  write(u, "(I0,1X,ES25.17E3)") n, e0
  do m = 1, n
    write(u, "(7(ES25.17E3,1X))") box(1)%XP(m, :), box(1)%FP(m, :), den(m)
  end do
  allocate(f_plain(n, 3))
  f_plain = box(1)%FP

  ! --- Maxwell velocities at 600 K on the device, temperature in K
  call Thermalizing_MC_DEV(box, ctrl, 600.0d0)
  call Cal_GlobalT_DEV(box, ctrl, tdev)
  call CopyOut_SimBox_DEV(box)
  ek = 0.0d0
  do m = 1, n
    ek = ek + 0.5d0 * box(1)%CM(1) * sum(box(1)%XP1(m, :)**2)
  end do
  thost = 2.0d0 * ek / (3.0d0 * n * CP_KB_CGS)
  write(u, "(A,2(1X,ES25.17E3))") "# temperature", tdev, thost

  ! --- an external force registered on the force class
  F0 = 0.01d0 * CP_EVERG / ANG
  call AddExtForce_ForceClass(fc, "PUSH", ForceArray=Push_ForceArray)
  call CalForce_ForceClass(box, ctrl, fc)
  call CopyOut_SimBox_DEV(box)
  maxd = maxval(abs(box(1)%FP(:, 1) - f_plain(:, 1) - F0))
  maxd = max(maxd, maxval(abs(box(1)%FP(:, 2:3) - f_plain(:, 2:3))))
  write(u, "(A,2(1X,ES25.17E3))") "# extforce", F0, maxd
  call Clear_ExtForce(fc, "PUSH")
  call CalForce_ForceClass(box, ctrl, fc)
  call CopyOut_SimBox_DEV(box)

  ! --- Predictor_DEV at ITIME 1 inside the damping window 1..1, step searched from HMX
  ctrl%DAMPTIME0 = 1
  ctrl%DAMPTIME1 = 1
  ctrl%IHDUP = -1
  ctrl%IT0 = 1
  ctrl%HMX = 2.0d-15
  ctrl%DMX = 0.002d-8
  write(u, "(A)") "# predictor-in"
  do m = 1, n
    write(u, "(6(ES25.17E3,1X))") box(1)%XP1(m, :), box(1)%FP(m, :)
  end do
  call Predictor_DEV(1, box, ctrl)
  call CopyOut_SimBox_DEV(box)
  write(u, "(A,1X,ES25.17E3)") "# predictor H", ctrl%H
  do m = 1, n
    write(u, "(6(ES25.17E3,1X))") box(1)%XP(m, :), box(1)%XP1(m, :)
  end do

  ! --- restart: 30 steps straight, against 20 steps + archive + restore on a new context + 10
  ctrl%H = 1.0d-15
  ctrl%IHDUP = 0
  ctrl%DAMPTIME1 = 0
  call Cal_NeighBoreList_DEV(box, ctrl)
  call CalForce_ForceClass(box, ctrl, fc)
  call Run_Steps_DEV(0, 20, box, ctrl)
  call CalEpot_ForceClass(box, ctrl, fc)
  call CopyOut_SimBox_DEV(box, WITHEPOT=.true.)
  stamp%AppType = "GMD"
  stamp%ITest = 1
  stamp%NBox = 1
  stamp%ITime = 20
  stamp%Time = 20.0d0 * ctrl%H
  call Archive_Config_SimBoxArray(box, stamp, trim(rstfile))
  call Run_Steps_DEV(20, 10, box, ctrl)
  call CopyOut_SimBox_DEV(box)
  allocate(ref(1))
  ref(1) = box(1)
  call Clear_Globle_Variables_DEV()

  stamp2%NBox = 1
  call Restore_Config_SimBoxArray(box2, stamp2, trim(rstfile))
  if (stamp2%ITime /= 20 .or. stamp2%AppType /= "GMD") stop 5
  call Initialize_Globle_Variables_DEV(box2, ctrl)
  call Init_Forcetable_Dev(box2, ctrl, fc)
  call CopyDISFrom_Host_to_Devices(box2)
  call Cal_NeighBoreList_DEV(box2, ctrl)
  call CalForce_ForceClass(box2, ctrl, fc)
  call Run_Steps_DEV(stamp2%ITime, 10, box2, ctrl)
  call CopyOut_SimBox_DEV(box2)
  dres(1) = maxval(abs(box2(1)%XP - ref(1)%XP))
  dres(2) = maxval(abs(box2(1)%XP1 - ref(1)%XP1))
  dres(3) = maxval(abs(box2(1)%DIS - ref(1)%DIS))
  dres(4) = maxval(abs(box2(1)%FP - ref(1)%FP))
  dres(5) = count(box2(1)%STATU /= ref(1)%STATU)
  write(u, "(A,5(1X,ES25.17E3))") "# restart", dres
  close(u)
  call Clear_Globle_Variables_DEV()
  write(*, "(A,I0,A,ES24.16)") "cgs_driver: n=", n, " E0(erg)=", e0
end program cgs_driver
