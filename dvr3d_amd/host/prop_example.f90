! Example PROP plug-in for the drop-in XPECT3: the property routines of the
! parity decks.  Parameters are read once from the file named by the
! environment variable DVR_PROP (default "prop.par"):
!   line 1: kind
!   kind 1 (synthetic, in the program's own coordinates r1, r2, x = cos theta;
!           u1 = r1 - 2.18, u2 = r2 - 3.20): five functions
!           1: 1   2: u1   3: u2   4: x   5: u1 u2 + 0.3 x^2
!           of which P(i) = f(IPROP(i)), i = 1 .. NPROPIN (COMMON /POT/)
!   kind 3 (tabulated):  line 2: the path of a table file, raw float64
!     (stream access): n, m, then n rows r1 r2 cos(theta) and m values --
!     e.g. another program's PROP evaluated on the grid of a job; PROP
!     returns the first NPROPIN values of the row at the point (a point
!     missing from the table stops the run).
! Users link their own SUBROUTINE PROP(P,R1,R2,XCOS,NPROPIN,NPRT) instead,
! exactly as for the reference program (xpect3.f90:12).
subroutine prop(p, r1, r2, xcos, npropin, nprt)
  implicit none
  integer, intent(in) :: npropin, nprt
  double precision, intent(out) :: p(npropin)
  double precision, intent(in) :: r1, r2, xcos
  integer :: iprop
  common /pot/ iprop(1000)
  integer, save :: kind = 0, ntab = 0, mtab = 0, last = 1
  double precision, allocatable, save :: tab(:, :)
  character(len=512) :: fname
  integer :: n, ios, i, j
  double precision :: ft(5), u1, u2
  if (kind == 0) then
    call get_environment_variable('DVR_PROP', fname, n, ios)
    if (ios /= 0 .or. n == 0) fname = 'prop.par'
    open(unit=94, file=trim(fname), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(6, *) 'prop: cannot open ', trim(fname)
      stop 1
    end if
    read(94, *) kind
    if (kind == 3) then
      read(94, '(a)') fname
      call load_table(trim(adjustl(fname)))
    else if (kind /= 1) then
      write(6, *) 'prop: unknown property kind', kind
      stop 1
    end if
    close(94)
  end if
  if (kind == 3) then
    if (npropin > mtab) then
      write(6, *) 'prop: the table holds fewer properties than requested'
      stop 1
    end if
    do j = 0, ntab - 1
      i = mod(last + j - 1, ntab) + 1
      if (abs(tab(1, i) - r1) + abs(tab(2, i) - r2) + abs(tab(3, i) - xcos) < 1.0d-12) then
        last = i + 1
        p = tab(4:3 + npropin, i)
        return
      end if
    end do
    write(6, *) 'prop: point missing from the table', r1, r2, xcos
    stop 1
  end if
  u1 = r1 - 2.18d0
  u2 = r2 - 3.20d0
  ft(1) = 1.0d0
  ft(2) = u1
  ft(3) = u2
  ft(4) = xcos
  ft(5) = u1 * u2 + 0.3d0 * xcos * xcos
  do i = 1, npropin
    if (iprop(i) < 1 .or. iprop(i) > 5) then
      write(6, *) 'prop: property number outside 1..5', iprop(i)
      stop 1
    end if
    p(i) = ft(iprop(i))
  end do
  if (nprt < 0) p = 0.0d0
contains

  subroutine load_table(path)
    character(len=*), intent(in) :: path
    double precision :: dn(2)
    integer :: u, ios
    open(newunit=u, file=path, access='stream', form='unformatted', status='old', &
         action='read', iostat=ios)
    if (ios /= 0) then
      write(6, *) 'prop: cannot open table ', path
      stop 1
    end if
    read(u) dn
    ntab = nint(dn(1))
    mtab = nint(dn(2))
    allocate(tab(3 + mtab, ntab))
    read(u) tab
    close(u)
  end subroutine

end subroutine prop
