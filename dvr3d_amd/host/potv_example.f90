! Example POTV plug-in for the drop-in executable: the analytic surfaces of
! the benchmark workloads (hartree, bohr).  Parameters are read once from the
! file named by the environment variable DVR_PES (default "pes.par"):
!   line 1: kind
!   kind 1 (AB2):  line 2: De a re        line 3: kb te(deg) c
!     V = De[(1-exp(-a(r1-re)))^2 + (1-exp(-a(r2-re)))^2]
!         + kb (cos(theta) - cos(te))^2 + c (r1-re)(r2-re)
!   kind 2 (ABC in Jacobi coordinates):  line 2: De a re1   line 3: kb re2 s
!     V = De[(1-exp(-a(r1-re1)))^2 + (1-exp(-a(d-re2)))^2] + kb (cos(theta)-1)^2
!         + s N(r1, r2, cos(theta))
!     d = sqrt(yy^2 + r2^2 + 2 yy r2 cos(theta)), yy = r1 (1 - g1): the A-B
!     distance (g1 from COMMON /MASS/, the pot_water.f convention for g2 = 0);
!     N is a standard normal number keyed by the IEEE bit patterns of the
!     three arguments (dvr_synth_normal, potv_noise.c).
!   kind 3 (tabulated):  line 2: the path of a table file, raw float64
!     (stream access): n, then n rows r1 r2 cos(theta) V -- e.g. another
!     program's POTV evaluated on the DVR grid of a job; POTV returns the V of
!     the row at the point (a point missing from the table stops the run).
! Users link their own SUBROUTINE POTV(V,R1,R2,XCOS) instead, exactly as for
! the reference program (docum.tex:31-45).
subroutine potv(v, r1, r2, xcos)
  use iso_c_binding, only: c_double
  implicit none
  interface
    real(c_double) function dvr_synth_normal(r1, r2, x) bind(C, name="dvr_synth_normal")
      import :: c_double
      real(c_double), value :: r1, r2, x
    end function
  end interface
  double precision, intent(out) :: v
  double precision, intent(in) :: r1, r2, xcos
  double precision :: xmass(3), g1, g2, xmassr(3)
  common /mass/ xmass, g1, g2, xmassr
  double precision, save :: p(6)
  integer, save :: kind = 0
  double precision, allocatable, save :: tab(:, :)
  integer, save :: ntab = 0, last = 1
  character(len=512) :: fname
  integer :: n, ios
  double precision :: t1, t2, te, yy, d
  if (kind == 0) then
    call get_environment_variable('DVR_PES', fname, n, ios)
    if (ios /= 0 .or. n == 0) fname = 'pes.par'
    open(unit=92, file=trim(fname), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(6, *) 'potv: cannot open ', trim(fname)
      stop 1
    end if
    read(92, *) kind
    if (kind == 3) then
      read(92, '(a)') fname
      close(92)
      call load_table(trim(adjustl(fname)))
    else
      read(92, *) p(1), p(2), p(3)
      read(92, *) p(4), te, p(6)
      close(92)
    end if
    if (kind == 3) then
      continue
    else if (kind == 1) then
      p(5) = cos(te * (4.0d0 * atan(1.0d0)) / 180.0d0)
    else if (kind == 2) then
      p(5) = te
    else
      write(6, *) 'potv: unknown PES kind', kind
      stop 1
    end if
  end if
  if (kind == 3) then
    v = lookup(r1, r2, xcos)
  else if (kind == 1) then
    t1 = 1.0d0 - exp(-p(2) * (r1 - p(3)))
    t2 = 1.0d0 - exp(-p(2) * (r2 - p(3)))
    v = p(1) * (t1 * t1 + t2 * t2) + p(4) * ((xcos - p(5)) * (xcos - p(5))) &
        + p(6) * ((r1 - p(3)) * (r2 - p(3)))
  else
    yy = r1 * (1.0d0 - g1)
    d = sqrt(yy * yy + r2 * r2 + 2.0d0 * yy * r2 * xcos)
    t1 = 1.0d0 - exp(-p(2) * (r1 - p(3)))
    t2 = 1.0d0 - exp(-p(2) * (d - p(5)))
    v = p(1) * (t1 * t1 + t2 * t2) + p(4) * ((xcos - 1.0d0) * (xcos - 1.0d0))
    if (p(6) /= 0.0d0) v = v + p(6) * dvr_synth_normal(r1, r2, xcos)
  end if
contains

  subroutine load_table(path)
    character(len=*), intent(in) :: path
    double precision :: dn
    integer :: u, ios
    open(newunit=u, file=path, access='stream', form='unformatted', status='old', &
         action='read', iostat=ios)
    if (ios /= 0) then
      write(6, *) 'potv: cannot open table ', path
      stop 1
    end if
    read(u) dn
    ntab = nint(dn)
    allocate(tab(4, ntab))
    read(u) tab
    close(u)
  end subroutine

  ! the row at (r1, r2, x), searched from the row after the previous hit
  ! (POTV is called in grid order, so the search is usually one step)
  double precision function lookup(r1, r2, x)
    double precision, intent(in) :: r1, r2, x
    integer :: i, j
    do j = 0, ntab - 1
      i = mod(last + j - 1, ntab) + 1
      if (abs(tab(1, i) - r1) + abs(tab(2, i) - r2) + abs(tab(3, i) - x) < 1.0d-12) then
        last = i + 1
        lookup = tab(4, i)
        return
      end if
    end do
    write(6, *) 'potv: point missing from the table', r1, r2, x
    stop 1
  end function

end subroutine potv
