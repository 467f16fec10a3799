! Evaluates a linked PROP(P,R1,R2,XCOS,NPROPIN,NPRT) on a list of points (test
! infrastructure: writes a property routine's values on a job's grid as
! data, so that the GPU parity tests need no property code).
!   stdin  (binary, stream access): n, npropin, nprt (int32), xmass(3), g1,
!          g2 (float64), zembed, zbisc (int32 logicals), iprop(npropin)
!          (int32), then n triples r1, r2, xcos (float64)
!   stdout (binary, stream access): n x npropin values (float64, the
!          npropin values of a point together)
program propgrid
  use iso_fortran_env, only: int32, real64
  implicit none
  double precision :: xmass(3), g1, g2
  logical :: zembed, zbisc
  common /mass/ xmass, g1, g2, zembed, zbisc
  integer :: iprop
  common /pot/ iprop(1000)
  integer(int32) :: n, npropin, nprt, lz(2)
  integer(int32), allocatable :: ip(:)
  integer :: i
  real(real64), allocatable :: p(:, :), v(:, :)
  open(unit=11, file='/dev/stdin', access='stream', form='unformatted', action='read')
  open(unit=12, file='/dev/stdout', access='stream', form='unformatted', action='write')
  read(11) n, npropin, nprt
  read(11) xmass, g1, g2
  read(11) lz
  zembed = lz(1) /= 0
  zbisc = lz(2) /= 0
  allocate(ip(npropin), p(3, n), v(npropin, n))
  read(11) ip
  iprop = 0
  iprop(1:npropin) = ip
  read(11) p
  do i = 1, n
    call prop(v(:, i), p(1, i), p(2, i), p(3, i), npropin, nprt)
  end do
  write(12) v
  close(12)
end program propgrid
