! Example DIPD plug-in for the drop-in DIPOLE3: the dipole surfaces of the
! parity decks (atomic units).  Parameters are read once from the file named
! by the environment variable DVR_DIP (default "dip.par"):
!   line 1: kind
!   kind 1 (synthetic, in the program's own coordinates r1, r2, x = cos theta;
!           u1 = r1 - 2.18, u2 = r2 - 3.20):
!     nu = 0:  mu_z = 0.35 + 0.12 u1 - 0.08 u2 + 0.21 x + 0.05 u1 x
!     nu = 1:  mu_x = sqrt(1 - x^2) (0.18 + 0.07 u1 + 0.04 u2 - 0.09 x)
!   kind 3 (tabulated):  line 2: the path of a table file, raw float64
!     (stream access): n, then n rows r1 r2 cos(theta) nu value -- e.g.
!     another program's DIPD evaluated on the integration grid of a job;
!     DIPD returns the value of the row at the point (a point missing from
!     the table stops the run).
! Users link their own SUBROUTINE DIPD(DIPC,R1,R2,XCOS,NU) instead, exactly
! as for the reference program (dipole3_segmented.f90:30-32).
subroutine dipd(dipc, r1, r2, xcos, nu)
  implicit none
  double precision, intent(out) :: dipc
  double precision, intent(in) :: r1, r2, xcos
  integer, intent(in) :: nu
  integer, save :: kind = 0
  double precision, allocatable, save :: tab(:, :)
  integer, save :: ntab = 0, last = 1
  character(len=512) :: fname
  integer :: n, ios
  double precision :: u1, u2, s
  if (kind == 0) then
    call get_environment_variable('DVR_DIP', fname, n, ios)
    if (ios /= 0 .or. n == 0) fname = 'dip.par'
    open(unit=93, file=trim(fname), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(6, *) 'dipd: cannot open ', trim(fname)
      stop 1
    end if
    read(93, *) kind
    if (kind == 3) then
      read(93, '(a)') fname
      call load_table(trim(adjustl(fname)))
    else if (kind /= 1) then
      write(6, *) 'dipd: unknown dipole kind', kind
      stop 1
    end if
    close(93)
  end if
  if (kind == 3) then
    dipc = lookup(r1, r2, xcos, nu)
  else
    u1 = r1 - 2.18d0
    u2 = r2 - 3.20d0
    if (nu == 0) then
      dipc = 0.35d0 + 0.12d0 * u1 - 0.08d0 * u2 + 0.21d0 * xcos + 0.05d0 * u1 * xcos
    else
      s = sqrt(max(0.0d0, 1.0d0 - xcos * xcos))
      dipc = s * (0.18d0 + 0.07d0 * u1 + 0.04d0 * u2 - 0.09d0 * xcos)
    end if
  end if
contains

  subroutine load_table(path)
    character(len=*), intent(in) :: path
    double precision :: dn
    integer :: u, ios
    open(newunit=u, file=path, access='stream', form='unformatted', status='old', &
         action='read', iostat=ios)
    if (ios /= 0) then
      write(6, *) 'dipd: cannot open table ', path
      stop 1
    end if
    read(u) dn
    ntab = nint(dn)
    allocate(tab(5, ntab))
    read(u) tab
    close(u)
  end subroutine

  ! the row at (r1, r2, x, nu), searched from the row after the previous
  ! hit (DIPD is called in grid order, so the search is usually one step)
  double precision function lookup(r1, r2, x, nu)
    double precision, intent(in) :: r1, r2, x
    integer, intent(in) :: nu
    integer :: i, j
    do j = 0, ntab - 1
      i = mod(last + j - 1, ntab) + 1
      if (abs(tab(1, i) - r1) + abs(tab(2, i) - r2) + abs(tab(3, i) - x) < 1.0d-12 .and. &
          nint(tab(4, i)) == nu) then
        last = i + 1
        lookup = tab(5, i)
        return
      end if
    end do
    write(6, *) 'dipd: point missing from the table', r1, r2, x, nu
    stop 1
  end function

end subroutine dipd
