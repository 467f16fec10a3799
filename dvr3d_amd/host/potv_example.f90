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
    read(92, *) p(1), p(2), p(3)
    read(92, *) p(4), te, p(6)
    close(92)
    if (kind == 1) then
      p(5) = cos(te * (4.0d0 * atan(1.0d0)) / 180.0d0)
    else if (kind == 2) then
      p(5) = te
    else
      write(6, *) 'potv: unknown PES kind', kind
      stop 1
    end if
  end if
  if (kind == 1) then
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
end subroutine potv
