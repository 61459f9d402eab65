!  w_fs_cgs.F90 -- the benchmark W Finnis-Sinclair pair and density functions in CGS, for the
!  Fortran callers that work in the reference's units (cgs_driver.F90, ref_glue_driver.F90): r in
!  cm, pair terms in erg and density in erg^2 (so that F = -sqrt(rho) is in erg), with the
!  NNFORCE/NEFORCE argument shape of MD_TypeDef_ForceTable.F90:46-78.

module W_FS_CGS
  use MD_ForceLib_HIP, only: KINDDF, CP_EVERG
  implicit none
  ! the benchmark W Finnis-Sinclair form (eV, Angstrom) evaluated for r in cm and returned in
  ! erg (pair) and erg^2 (density), so that F = -sqrt(rho) is in erg
  real(KINDDF), parameter :: ANG = 1.0d-8
  real(KINDDF), parameter :: C = 3.25d0, C0 = 47.1346d0, C1 = -33.7666d0, C2 = 6.2542d0
  real(KINDDF), parameter :: A = 1.8964d0, D = 4.4002d0
contains
  subroutine W_NN_CGS(R, POTR, FPOTR)
    real(KINDDF), intent(in) :: R
    real(KINDDF), intent(out) :: POTR, FPOTR
    real(KINDDF) :: ra, rc, poly
    ra = R / ANG
    if (ra <= C) then
      rc = ra - C
      poly = C0 + C1 * ra + C2 * ra * ra
      POTR = 0.5d0 * (rc * rc * poly) * CP_EVERG
      FPOTR = -(2.0d0 * rc * poly + rc * rc * (C1 + 2.0d0 * C2 * ra)) * CP_EVERG / ANG
    else
      POTR = 0.0d0
      FPOTR = 0.0d0
    end if
  end subroutine W_NN_CGS
  subroutine W_NE_CGS(R, POTB, FPOTB)
    real(KINDDF), intent(in) :: R
    real(KINDDF), intent(out) :: POTB, FPOTB
    real(KINDDF) :: ra, rd
    ra = R / ANG
    if (ra <= D) then
      rd = ra - D
      POTB = A * A * rd * rd * CP_EVERG * CP_EVERG
      FPOTB = -(A * A * 2.0d0 * rd) * CP_EVERG * CP_EVERG / ANG
    else
      POTB = 0.0d0
      FPOTB = 0.0d0
    end if
  end subroutine W_NE_CGS
end module W_FS_CGS
