!  fs_w_driver.F90 -- a Fortran caller of the drop-in layer: BCC W box (NC^3 conventional cells,
!  a = 3.165 A), Finnis-Sinclair W tables built from Fortran potential routines through
!  Register_ForceTableProc / Create_Interaction_ForceTable, then neighbour list, force and energy
!  on the device, NSTEPS velocity-Verlet steps (list rebuilt every 10), and the same quantities at
!  the end.  Units: eV, Angstrom, fs, masses in eV fs^2 / A^2.
!
!  usage: fs_w_driver OUTFILE [NSTEPS [NC]]
!  OUTFILE: line 1 "N E0 0"; N lines "x y z fx fy fz" of the start; N lines of the end; a line
!  "# energies E0 E1" (total potential energy at the start and at the end), a line
!  "# stress_sum S11 S12 ... S33" (per-atom stresses of Cal_AtomicStress summed over the atoms at
!  the end), a line "# temperature T" (Cal_GlobalT_DEV after VelScaling_DEV to 300 K), a line
!  "# steepest" and N lines "x y z" after Do_Steepest_Forsteps_DEV (list rebuilt on the end
!  positions, STEEPEST_MiDelE = 1e-6 eV, moves bounded by RR = a), then a line
!  "# copy-in X T": x of atom 1 moved by 0.01 A and velocities zeroed on the host, copied in with
!  CopyXPFrom_Host_to_Devices / CopyXP1From_Host_to_Devices and read back (X, and Cal_GlobalT_DEV).

module W_FS_Potential
  use MD_ForceLib_HIP, only: KINDDF
  implicit none
  ! pair (r-c)^2 (c0 + c1 r + c2 r^2) for r <= c; rho = A^2 (r-d)^2 for r <= d; F = -sqrt(rho)
  real(KINDDF), parameter :: C = 3.25d0, C0 = 47.1346d0, C1 = -33.7666d0, C2 = 6.2542d0
  real(KINDDF), parameter :: A = 1.8964d0, D = 4.4002d0
contains
  ! NN routine: POTR = V/2, FPOTR = -dV/dr
  subroutine W_NN(R, POTR, FPOTR)
    real(KINDDF), intent(in) :: R
    real(KINDDF), intent(out) :: POTR, FPOTR
    real(KINDDF) :: rc, poly
    if (R <= C) then
      rc = R - C
      poly = C0 + C1 * R + C2 * R * R
      POTR = 0.5d0 * (rc * rc * poly)
      FPOTR = -(2.0d0 * rc * poly + rc * rc * (C1 + 2.0d0 * C2 * R))
    else
      POTR = 0.0d0
      FPOTR = 0.0d0
    end if
  end subroutine W_NN
  ! NE routine: POTB = rho, FPOTB = -drho/dr
  subroutine W_NE(R, POTB, FPOTB)
    real(KINDDF), intent(in) :: R
    real(KINDDF), intent(out) :: POTB, FPOTB
    real(KINDDF) :: rd
    if (R <= D) then
      rd = R - D
      POTB = A * A * rd * rd
      FPOTB = -(A * A * 2.0d0 * rd)
    else
      POTB = 0.0d0
      FPOTB = 0.0d0
    end if
  end subroutine W_NE
  subroutine W_EMB(RHO, FRHO, DFRHO)
    real(KINDDF), intent(in) :: RHO
    real(KINDDF), intent(out) :: FRHO, DFRHO
    if (RHO > 0.0d0) then
      FRHO = -sqrt(RHO)
      DFRHO = -0.5d0 / sqrt(RHO)
    else
      FRHO = 0.0d0
      DFRHO = 0.0d0
    end if
  end subroutine W_EMB
end module W_FS_Potential

program fs_w_driver
  use iso_c_binding
  use MSM_HIP_C, only: msm_drand32, msm_thermalize_mc
  use MD_ForceLib_HIP
  use W_FS_Potential
  implicit none
  real(KINDDF), parameter :: AW = 3.165d0, SIGMA = 0.03d0, TEMP = 300.0d0
  real(KINDDF), parameter :: KB = 8.617333262d-5
  real(KINDDF), parameter :: AMU = 1.66053906660d-27 * 1.0d10 / 1.602176634d-19
  real(KINDDF), parameter :: TWOPI = 6.283185307179586d0
  type(SimMDBox) :: box(1)
  type(SimMDCtrl) :: ctrl
  type(MDForceClassGPU) :: fc
  character(len=512) :: outfile, arg
  integer :: nsteps, nc, n, i, j, k, l, m, u
  integer(c_int) :: seed(2), rc
  real(KINDDF) :: e0, e1, z1, z2, x
  real(KINDDF), allocatable :: avp(:, :)
  real(KINDDF) :: curt

  if (command_argument_count() < 1) then
    write(*, "(A)") "usage: fs_w_driver OUTFILE [NSTEPS [NC]]"
    stop 2
  end if
  call get_command_argument(1, outfile)
  nsteps = 100
  nc = 16
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) nsteps
  end if
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) nc
  end if

  ! --- box: BCC sites of nc^3 cells in [-L/2, L/2), Gaussian displacements from DRAND32
  n = 2 * nc * nc * nc
  box(1)%NPRT = n
  box(1)%NGROUP = 1
  box(1)%ZL = nc * AW
  box(1)%RR = AW
  box(1)%BOXLOW = -0.5d0 * box(1)%ZL
  allocate(box(1)%XP(n, 3), box(1)%XP1(n, 3), box(1)%ITYP(n), box(1)%STATU(n), box(1)%CM(1))
  allocate(box(1)%PTYPE(1, 1))
  box(1)%ITYP = 1
  box(1)%STATU = 1
  box(1)%CM(1) = 183.84d0 * AMU
  box(1)%PTYPE(1, 1) = 1
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
  seed = [int(4242, c_int), int(777, c_int)]
  do k = 1, 3
    do m = 1, n
      z1 = msm_drand32(seed)
      z2 = msm_drand32(seed)
      x = box(1)%XP(m, k) + SIGMA * sqrt(-2.0d0 * log(max(z1, 1.0d-300))) * cos(TWOPI * z2)
      ! periodic wrap into [BOXLOW, BOXLOW + ZL)
      x = box(1)%BOXLOW(k) + modulo(x - box(1)%BOXLOW(k), box(1)%ZL(k))
      if (x >= box(1)%BOXLOW(k) + box(1)%ZL(k)) x = x - box(1)%ZL(k)
      box(1)%XP(m, k) = x
    end do
  end do
  seed = [int(1001, c_int), int(2002, c_int)]
  rc = msm_thermalize_mc(int(n, c_int), box(1)%ITYP, box(1)%CM, TEMP, KB, seed, box(1)%XP1)
  if (rc /= 0) stop 3

  ! --- control block and tables
  allocate(ctrl%RU(1, 1), ctrl%NB_RM(1, 1))
  ctrl%RU = 4.4002d0
  ctrl%NB_RM = 4.7002d0
  ctrl%NB_UPTAB = 10
  ctrl%NB_MXNBS = 64
  ctrl%NUMFTABR = 5000
  ctrl%NUMFTABE = 5000
  ctrl%H = 1.0d0
  ctrl%IFPD = 1
  call Register_ForceClass("FS_TYPE", fc)
  call Register_ForceTableProc(1, fc%ForceTable, W_NN, W_NE, W_EMB, NOTE="W-W Finnis-Sinclair")
  call Create_Interaction_ForceTable(box(1), ctrl, fc%ForceTable)

  ! --- device: list, force, energy at the start
  call Initialize_Globle_Variables_DEV(box, ctrl)
  call Init_Forcetable_Dev(box, ctrl, fc)
  call Cal_NeighBoreList_DEV(box, ctrl)
  call CalForce_ForceClass(box, ctrl, fc)
  call CalEpot_ForceClass(box, ctrl, fc)
  call CopyOut_SimBox_DEV(box, WITHEPOT=.true.)
  e0 = sum(box(1)%EPOT)
  open(newunit=u, file=trim(outfile), status="replace", action="write")
  write(u, "(I0,2(1X,ES25.17E3))") n, e0, 0.0d0
  do m = 1, n
    write(u, "(6(ES25.17E3,1X))") box(1)%XP(m, :), box(1)%FP(m, :)
  end do

  ! --- NSTEPS steps: the first one through the per-step routine, the rest in one fused run
  if (nsteps > 0) then
    call For_One_Step(1, box, ctrl, fc)
    if (nsteps > 1) call Run_Steps_DEV(1, nsteps - 1, box, ctrl)
  end if
  call CalEpot_ForceClass(box, ctrl, fc)
  call CopyOut_SimBox_DEV(box, WITHEPOT=.true.)
  e1 = sum(box(1)%EPOT)
  do m = 1, n
    write(u, "(6(ES25.17E3,1X))") box(1)%XP(m, :), box(1)%FP(m, :)
  end do
  write(u, "(A,2(1X,ES25.17E3))") "# energies", e0, e1
  allocate(avp(n, 9))
  call Cal_AtomicStress(box, ctrl, avp)
  write(u, "(A,9(1X,ES25.17E3))") "# stress_sum", (sum(avp(:, m)), m = 1, 9)
  ! thermalisation and quench through this layer: scale to 300 K and read the temperature back,
  ! then steepest descent on a list of the current positions, final positions written
  call VelScaling_DEV(box, ctrl, 300.0d0)
  call Cal_GlobalT_DEV(box, ctrl, curt)
  write(u, "(A,1X,ES25.17E3)") "# temperature", curt
  call Cal_NeighBoreList_DEV(box, ctrl)
  ctrl%STEEPEST_MiDelE = 1.0d-6
  call Do_Steepest_Forsteps_DEV(box, ctrl, fc, 200, CP_DAMPSCHEME_ST)
  call CopyOut_SimBox_DEV(box, WITHEPOT=.false.)
  write(u, "(A)") "# steepest"
  do m = 1, n
    write(u, "(3(ES25.17E3,1X))") box(1)%XP(m, :)
  end do
  box(1)%XP(1, 1) = box(1)%XP(1, 1) + 0.01d0
  box(1)%XP1 = 0.0d0
  call CopyXPFrom_Host_to_Devices(box)
  call CopyXP1From_Host_to_Devices(box)
  call Cal_GlobalT_DEV(box, ctrl, curt)
  call CopyOut_SimBox_DEV(box, WITHEPOT=.false.)
  write(u, "(A,2(1X,ES25.17E3))") "# copy-in", box(1)%XP(1, 1), curt
  close(u)
  call Clear_Globle_Variables_DEV()
  write(*, "(A,I0,A,ES24.16,A,ES24.16)") "fs_w_driver: n=", n, " E0=", e0, " E1=", e1
end program fs_w_driver
