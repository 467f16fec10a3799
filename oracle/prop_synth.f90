! Synthetic property routine of the XPECT3 parity decks (test
! infrastructure: linked into the reference XPECT3 build by oracle/Makefile,
! target xpect).  Five smooth functions of the program's own coordinates
! (u1 = r1 - 2.18, u2 = r2 - 3.20, x = cos theta):
!   1: 1   2: u1   3: u2   4: x   5: u1 u2 + 0.3 x^2
! Property 1 makes every expectation value the norm of the wavefunction.
! COMMON /POT/ IPROP selects which of the five are returned, as XPECT3
! passes them.  The same formula is coded independently by the drop-in's
! example property routine (dvr3d_amd/host/prop_example.f90, kind 1).
subroutine prop(p, r1, r2, xcos, npropin, nprt)
  implicit none
  integer, intent(in) :: npropin, nprt
  double precision, intent(out) :: p(npropin)
  double precision, intent(in) :: r1, r2, xcos
  integer :: iprop
  common /pot/ iprop(1000)
  double precision :: ft(5), u1, u2
  integer :: i
  u1 = r1 - 2.18d0
  u2 = r2 - 3.20d0
  ft(1) = 1.0d0
  ft(2) = u1
  ft(3) = u2
  ft(4) = xcos
  ft(5) = u1 * u2 + 0.3d0 * xcos * xcos
  do i = 1, npropin
    p(i) = ft(iprop(i))
  end do
end subroutine prop
