! Example POTV plug-in for the drop-in executable: the analytic AB2 surface
! of the benchmark workloads,
!   V = De[(1-exp(-a(r1-re)))^2 + (1-exp(-a(r2-re)))^2]
!       + kb (cos(theta) - cos(te))^2 + c (r1-re)(r2-re)
! (hartree, bohr).  Parameters are read once from the file named by the
! environment variable DVR_PES (default "pes.par"):
!   line 1: 1            line 2: De a re            line 3: kb te(deg) c
! Users link their own SUBROUTINE POTV(V,R1,R2,XCOS) instead, exactly as for
! the reference program (docum.tex:31-45).
subroutine potv(v, r1, r2, xcos)
  implicit none
  double precision, intent(out) :: v
  double precision, intent(in) :: r1, r2, xcos
  double precision :: xmass(3), g1, g2, xmassr(3)
  common /mass/ xmass, g1, g2, xmassr
  double precision, save :: p(6)
  logical, save :: have = .false.
  character(len=512) :: fname
  integer :: n, ios, kind
  double precision :: t1, t2, te
  if (.not. have) then
    call get_environment_variable('DVR_PES', fname, n, ios)
    if (ios /= 0 .or. n == 0) fname = 'pes.par'
    open(unit=92, file=trim(fname), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(6, *) 'potv: cannot open ', trim(fname)
      stop 1
    end if
    read(92, *) kind
    read(92, *) p(1), p(2), p(3)
    read(92, *) p(4), te, p(6)
    close(92)
    if (kind /= 1) then
      write(6, *) 'potv: only the AB2 form (kind 1) is provided'
      stop 1
    end if
    p(5) = cos(te * (4.0d0 * atan(1.0d0)) / 180.0d0)
    have = .true.
  end if
  t1 = 1.0d0 - exp(-p(2) * (r1 - p(3)))
  t2 = 1.0d0 - exp(-p(2) * (r2 - p(3)))
  v = p(1) * (t1 * t1 + t2 * t2) + p(4) * ((xcos - p(5)) * (xcos - p(5))) + p(6) * ((r1 - p(3)) * (r2 - p(3)))
end subroutine potv
