! Synthetic dipole surface of the DIPOLE3 parity decks (test infrastructure:
! linked into the reference DIPOLE3 build by oracle/Makefile, target dip).
! The components are smooth functions of the program's own coordinates
! (r1, r2, cos theta) with no coordinate transformation, so the test
! exercises DIPOLE3's quadrature, angular transformation and block algebra
! only.  The same formula is coded independently by the drop-in's example
! surface (dvr3d_amd/host/dipd_example.f90, kind 1).
!   nu = 0 (component along the embedding axis):
!     mu_z = 0.35 + 0.12 u1 - 0.08 u2 + 0.21 x + 0.05 u1 x
!   nu = 1 (perpendicular component):
!     mu_x = sqrt(1 - x^2) (0.18 + 0.07 u1 + 0.04 u2 - 0.09 x)
!   with u1 = r1 - 2.18, u2 = r2 - 3.20, x = cos theta (atomic units).
subroutine dipd(dipc, r1, r2, xcos, nu)
  implicit none
  double precision, intent(out) :: dipc
  double precision, intent(in) :: r1, r2, xcos
  integer, intent(in) :: nu
  double precision :: u1, u2, s
  u1 = r1 - 2.18d0
  u2 = r2 - 3.20d0
  if (nu == 0) then
    dipc = 0.35d0 + 0.12d0 * u1 - 0.08d0 * u2 + 0.21d0 * xcos + 0.05d0 * u1 * xcos
  else
    s = sqrt(max(0.0d0, 1.0d0 - xcos * xcos))
    dipc = s * (0.18d0 + 0.07d0 * u1 + 0.04d0 * u2 - 0.09d0 * xcos)
  end if
end subroutine dipd
