! Evaluates a linked POTV(V,R1,R2,XCOS) on a list of points (test
! infrastructure: writes the potential of a reference job's DVR grid as data,
! so that the GPU parity tests need no PES code).
!   stdin  (binary, stream access): n (int32), xmass(3), g1, g2, xmassr(3)
!          (float64), then n triples r1, r2, xcos (float64)
!   stdout (binary, stream access): n values V (float64)
program potgrid
  use iso_fortran_env, only: int32, real64, input_unit, output_unit
  implicit none
  double precision :: xmass(3), g1, g2, xmassr(3)
  common /mass/ xmass, g1, g2, xmassr
  integer(int32) :: n
  integer :: i
  real(real64), allocatable :: p(:, :), v(:)
  open(unit=11, file='/dev/stdin', access='stream', form='unformatted', action='read')
  open(unit=12, file='/dev/stdout', access='stream', form='unformatted', action='write')
  read(11) n
  read(11) xmass, g1, g2, xmassr
  allocate(p(3, n), v(n))
  read(11) p
  do i = 1, n
    call potv(v(i), p(1, i), p(2, i), p(3, i))
  end do
  write(12) v
  close(12)
end program potgrid
