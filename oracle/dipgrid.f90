! Evaluates a linked DIPD(DIPC,R1,R2,XCOS,NU) on a list of points (test
! infrastructure: writes the dipole surface of a reference job's integration
! grid as data, so that the GPU parity tests need no surface code).
!   stdin  (binary, stream access): n (int32), xmass(3), g1, g2 (float64),
!          zembed, zbisc (int32 logicals), then n records r1, r2, xcos
!          (float64) and nu (float64, integral)
!   stdout (binary, stream access): n values DIPC (float64)
program dipgrid
  use iso_fortran_env, only: int32, real64
  implicit none
  double precision :: xmass(3), g1, g2
  logical :: zembed, zbisc
  common /mass/ xmass, g1, g2, zembed, zbisc
  logical :: zmors1, znco1, znco2, zprint, zpmin, ztra, zstart, zmors2
  common /logic/ zmors1, znco1, znco2, zprint, zpmin, ztra, zstart, zmors2
  integer(int32) :: n, lz(2)
  integer :: i
  real(real64), allocatable :: p(:, :), v(:)
  open(unit=11, file='/dev/stdin', access='stream', form='unformatted', action='read')
  open(unit=12, file='/dev/stdout', access='stream', form='unformatted', action='write')
  read(11) n
  read(11) xmass, g1, g2
  read(11) lz
  zembed = lz(1) /= 0
  zbisc = lz(2) /= 0
  zprint = .false.
  allocate(p(4, n), v(n))
  read(11) p
  do i = 1, n
    call dipd(v(i), p(1, i), p(2, i), p(3, i), nint(p(4, i)))
  end do
  write(12) v
  close(12)
end program dipgrid
