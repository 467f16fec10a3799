! Drop-in DIPOLE3 on the MI355X: line strengths between the vib-rot
! wavefunctions of two runs (source/dipole3_segmented.f90, cited as P:<line>).
! Same namelist PRT and cards on stdin, the same ket/bra input files (the
! KVEC file of ROTLEV3 or the IWAVE file of a J = 0 DVR3DRJZ run), the same
! stdout table and SPECTRA file (ITRA); the wavefunctions' evaluation at the
! angular integration points and the transition-moment sums run on the GPU
! through dvr_dipole3 (libdvr3d_amd).
!
! The host does what the reference's host does: INSIZE's checks (P:198-620),
! GENIND (P:656-704), the Gauss-Legendre points and the dipole surface on the
! grid (LAGPT P:1141-1230), the associated Legendre values (ASLEG P:1234-1308)
! and JTRAN's transformation matrix (P:1564-1645), the order and factors of
! the k -> k' blocks (DMAIN P:708-1115, THREEJ P:1778-1822), the restart file
! (RDSCR/WRSCR P:2022-2041) and SPECT/OUTPT (P:1825-2002).  The dipole surface is the
! user's SUBROUTINE DIPD(DIPC,R1,R2,XCOS,NU), linked as for the reference.
! Not supported (refused with a message): IDIA = -2 files (symmetrised Radau,
! the ROTLEV3B output) and pairs of non-Coriolis files.
program dipole3_amd
  use iso_c_binding
  use dvr3d_amd_api
  implicit none

  ! namelist PRT (P:145-146) with the block data defaults (P:192-194); the
  ! logical switches live in COMMON /LOGIC/ as in the reference, where a
  ! user's DIPD may declare them
  logical :: zmors1, znco1, znco2, zprint, zpmin, ztra, zstart, zmors2
  common /logic/ zmors1, znco1, znco2, zprint, zpmin, ztra, zstart, zmors2
  double precision :: xmass(3), g1, g2
  logical :: zembed, zbisc
  common /mass/ xmass, g1, g2, zembed, zbisc
  integer :: iket, ibra, itra, iscr, ires, nblock
  namelist /prt/ zprint, zpmin, ztra, zstart, iket, ibra, itra, iscr, ires, nblock

  type wavefile
    integer :: unit, idia, ipar, lmax, npnt1, npnt2, jrot, kmin, neval, jk, mbass, mbass0
    logical :: zemb, zmor1, zmor2, znco
    double precision :: xm(3), g1, g2, morse(6)
    integer, allocatable :: lmin(:), lbass(:), nbass(:)
    double precision, allocatable :: r1(:), r2(:), e(:)
    ! J = 0 IWAVE angular records (D:973-977)
    integer :: kz, maxleg, nidvr
    double precision, allocatable :: pleg(:, :)
  end type

  type blockdata
    integer :: nin = 0
    double precision, allocatable :: vec(:, :), ang(:, :)
  end type

  double precision, parameter :: x0 = 0.0d0, x1 = 1.0d0, x2 = 2.0d0
  double precision, parameter :: autocm = 2.19474624d+05, autode = 2.5417662d0, &
                                 detosec = 3.136186d-07
  character(len=8) :: title(9)
  type(wavefile) :: ket, bra
  integer :: idia, ipar1, ipar2, jrot1, jrot2, kmin1, kmin2, jk1, jk2, mblock
  integer :: npot, nv1, nv2, ibase1, ibase2, ne1, ne2, nn2, ipot, jdia, nrad
  double precision, allocatable :: xd(:), wtd(:), dipol(:, :), tz(:, :), tx(:, :)
  type(blockdata), allocatable, target :: kblk(:), bblk(:)
  integer, allocatable :: pb1(:), pb2(:), pcomp(:)
  double precision, allocatable :: pfac(:), binom(:, :)
  integer :: npair, nbin

  zprint = .false.; zpmin = .false.; ztra = .true.; zstart = .false.
  zmors1 = .true.; zmors2 = .true.
  iket = 11; ibra = 12; itra = 13; iscr = 24; ires = 0; nblock = 1000

  write(6, 200)
200 format(//,5x,'Program DIPOLE3 (MI355X / gfx950 build of the line-strength step):',/)
  read(5, prt)
  read(5, '(9a8)') title
  write(6, '(5x,9a8)') title
  call insize()
  call dmain()

contains

  ! the file a unit without an OPEN statement is connected to (fort.<unit>)
  character(len=16) function unit_file(iu)
    integer, intent(in) :: iu
    write(unit_file, '(a,i0)') 'fort.', iu
  end function

  subroutine fatal(msg)
    character(len=*), intent(in) :: msg
    write(6, '(//,5x,a,/)') msg
    stop 1
  end subroutine

  subroutine open_wave(w, iu)
    type(wavefile), intent(inout) :: w
    integer, intent(in) :: iu
    integer :: ios
    w%unit = iu
    open(unit=iu, form='unformatted', status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(6, '(/5x,a,i4)') 'cannot open wavefunction file on stream ', iu
      stop 1
    end if
    read(iu) w%idia, w%ipar, w%lmax, w%npnt1, w%npnt2, w%jrot, w%kmin, w%neval
    read(iu) w%zemb, w%zmor1, w%zmor2, w%xm, w%g1, w%g2, w%znco
  end subroutine

  ! INSIZE (P:198-620): the checks that bra and ket belong together, the
  ! block counts and the cards of the integration
  subroutine insize()
    double precision, parameter :: toler = 1.0d-3
    integer :: i
    write(6, 1057) iket, ibra, itra
1057 format(//5x,'input ket wavefunctions read from stream iket  =',i4, &
            /5x,'input bra wavefunctions read from stream ibra  =',i4, &
            /5x,'output transition data written to stream itra  =',i4)
    if (ires <= 0) write(6, '(5x,a,i4)') 'restart data written to           stream iscr  =', iscr
    if (ires > 0) write(6, '(5x,a,i4)') 'RESTART RUN: input data read from stream iscr  =', iscr
    call open_wave(ket, iket)
    call open_wave(bra, ibra)
    if (ket%idia /= bra%idia) then
      write(6, 998) ket%idia, bra%idia
998   format(//,5x,'** fatal ** diatomic mismatch',/5x,'idia1=',i2,'  idia2=',i2,/)
      stop 1
    end if
    idia = ket%idia
    write(6, '(/,5x,a,i4)') 'diatomic parameter idia  =', idia
    if (idia <= -2) call fatal('IDIA = -2 (symmetrised Radau) wavefunctions are not supported')
    ipar1 = ket%ipar
    ipar2 = bra%ipar
    jrot1 = ket%jrot
    jrot2 = bra%jrot
    kmin1 = ket%kmin
    kmin2 = bra%kmin
    znco1 = ket%znco
    znco2 = bra%znco
    if (jrot1 == 0 .or. jrot2 == 0) then
      zembed = ket%zemb
      if (jrot1 == 0) zembed = bra%zemb
      if (jrot1 == 0 .and. jrot2 == 0) call fatal('j = 0 -> 0 not allowed: stop')
    else
      if (ket%zemb .neqv. bra%zemb) call fatal('** fatal ** embedding mismatch')
      zembed = ket%zemb
    end if
    if (zembed) then
      if (ipar1 /= ipar2 .and. idia == 2) call parity_fatal()
    else
      if (ipar1 == ipar2 .and. idia == 2) call parity_fatal()
    end if
    zbisc = .false.
    if (ket%zmor1 .neqv. bra%zmor1) call fatal('** fatal ** r1 radial function mismatch')
    zmors1 = ket%zmor1
    if (ket%zmor2 .neqv. bra%zmor2) call fatal('** fatal ** r2 radial function mismatch')
    zmors2 = bra%zmor2
    write(6, '(/,5x,a,i4)') 'number of co-ordinates   =', 3
    if (zembed) then
      write(6, 886) 2
    else
      write(6, 886) 1
    end if
886 format(/,5x,'z axis embedded along r',i1,' co-ordinate'/)
    call radial_kind(1, zmors1)
    if (ket%npnt1 /= bra%npnt1) call fatal('** fatal ** r1 radial function mismatch (dvr points)')
    call radial_kind(2, zmors2)
    if (ket%npnt2 /= bra%npnt2) call fatal('** fatal ** r2 radial function mismatch (dvr points)')
    if (zpmin) write(6, '(/5x,a)') ' Minimum printing requested'
    if (abs(bra%g1 - ket%g1) > toler .or. abs(bra%g2 - ket%g2) > toler) &
      call fatal('co-ordinate system incompatible')
    g1 = ket%g1
    g2 = bra%g2
    do i = 1, 3
      if (abs(bra%xm(i) - ket%xm(i)) > toler) call fatal('masses incompatible')
      xmass(i) = ket%xm(i)
    end do
    read(bra%unit) bra%morse
    read(ket%unit) ket%morse
    do i = 1, 6
      if (abs(ket%morse(i) - bra%morse(i)) > toler) &
        call fatal('radial basis parameters (re, diss, we) incompatible')
    end do
    ! non-Coriolis files next to coupled ones (P:485-503)
    if (znco1 .and. znco2) call fatal('non-Coriolis wavefunction pairs are not supported')
    if (znco1 .and. .not. znco2) then
      if (abs(jrot1) > 1) call fatal('attempt to mix no coriolis and coupled calcs: stop')
      kmin1 = 0
    end if
    if (znco2 .and. .not. znco1) then
      if (abs(jrot2) > 1) call fatal('attempt to mix no coriolis and coupled calcs: stop')
      kmin2 = 0
    end if
    if (jrot1 == 0) kmin1 = 1
    if (jrot2 == 0) kmin2 = 1
    jk1 = jrot1 + kmin1
    jk2 = jrot2 + kmin2
    mblock = jk1 + jk2 + min(jk1, jk2) - 2
    nblock = min(nblock, mblock)
    if (nblock < mblock) write(6, '(/i7,a,i4)') nblock, ' blocks to be calculated out a maximum of', &
                                                 mblock
    ket%jk = jk1
    bra%jk = jk2
    nrad = ket%npnt1 * ket%npnt2
    ket%mbass = nrad * jk1 * ket%lmax
    bra%mbass = nrad * jk2 * bra%lmax
    read(5, '(5i5)') npot, nv1, nv2, ibase1, ibase2
    if (ibase1 <= 0 .or. ibase1 > ket%neval) ibase1 = 1
    if (ibase2 <= 0 .or. ibase2 > bra%neval) ibase2 = 1
    write(6, '(/,9x,a,/)') 'parameters passed to dipole for the ket & the bra'
    write(6, 201) ket%mbass, bra%mbass, ket%npnt1, bra%npnt1, ket%npnt2, bra%npnt2, &
                  ket%lmax, bra%lmax, jrot1, jrot2, ipar1, ipar2, kmin1, kmin2, &
                  ket%neval, bra%neval, nv1, nv2, ibase1, ibase2
201 format(5x,'total number of basis functions    = ',i6,3x,i6,/ &
           5x,'number of r1 radial dvr points     = ',i6,3x,i6,/ &
           5x,'number of r2 radial dvr points     = ',i6,3x,i6,/ &
           5x,'number of angular functions        = ',i6,3x,i6,/ &
           5x,'j, total angular momentum          = ',i6,3x,i6,/ &
           5x,'ipar, vibrational parity           = ',i6,3x,i6,/ &
           5x,'kmin= 1-p, where j + p is parity,  = ',i6,3x,i6,/ &
           5x,'total vib-rot functions available  = ',i6,3x,i6,/ &
           5x,'vib-rot functions taken (0 = all)  = ',i6,3x,i6,/ &
           5x,'neva1 and neval2 used (all)        = ',i6,3x,i6,/)
    if (mod(jrot1 + jrot2 + 1 + kmin1 + kmin2, 2) /= 0 .or. abs(jrot1 - jrot2) > 1) &
      call fatal('selection rules violated')
    if (nv1 > 0) ne1 = min(nv1, ket%neval - ibase1 + 1)
    if (nv1 <= 0) ne1 = ket%neval - ibase1 + 1
    if (nv2 > 0) ne2 = min(nv2, bra%neval - ibase2 + 1)
    if (nv2 <= 0) ne2 = bra%neval - ibase2 + 1
    jdia = max(1, idia)
    if (mod(npot, 2) /= 0) npot = npot + 1
    nn2 = npot / 2
    ipot = npot
    if (idia >= 2) ipot = npot / 2
  end subroutine insize

  subroutine parity_fatal()
    write(6, 997) ipar1, ipar2
997 format(//,5x,'** fatal ** parity mismatch, spin forbidden',/,5x,'ipar1=',i2,'  ipar2=',i2,/)
    stop 1
  end subroutine

  subroutine radial_kind(ir, zm)
    integer, intent(in) :: ir
    logical, intent(in) :: zm
    if (zm) then
      write(6, '(5x,a,i1,a)') 'morse functions in r', ir, ' radial basis set'
    else
      write(6, '(5x,a,i1,a)') 'spherical functions in r', ir, ' radial basis set'
    end if
  end subroutine

  ! GENIND (P:656-704) and the rest of a file: the radial grids, energies
  ! and the vectors of levels ibase .. ibase + ne - 1 of every k block
  subroutine read_body(w, ne, ibase, jay_ipar, blk)
    type(wavefile), intent(inout) :: w
    integer, intent(in) :: ne, ibase, jay_ipar
    type(blockdata), allocatable, intent(out) :: blk(:)
    integer :: k, i, meval, nb, jstart, jj0, nang, ios
    double precision :: dum
    double precision, allocatable :: xalf(:)
    allocate(w%lmin(w%jk), w%lbass(w%jk), w%nbass(w%jk), blk(w%jk))
    read(w%unit, iostat=ios) w%mbass0, w%lmin, w%lbass, w%nbass
    if (ios /= 0 .or. w%mbass0 > w%mbass) then
      write(6, 220) w%unit, w%mbass, w%mbass0
220   format(//,5x,'basis function dimensions in error',/,5x,'unit =',i3,' mbass =',i7, &
             ' expected, =',i7,' found')
      stop 1
    end if
    allocate(w%r1(w%npnt1), w%r2(w%npnt2))
    read(w%unit) w%r1
    read(w%unit) w%r2
    if (w%jk <= 1) then
      ! J = 0 IWAVE: angular records, then the wavefunctions (D:3733-3841)
      allocate(xalf(w%lmax))
      read(w%unit) xalf
      read(w%unit) w%kz, w%maxleg, w%nidvr
      allocate(w%pleg(0:w%maxleg, w%nidvr))
      read(w%unit) w%pleg
      read(w%unit) meval
      allocate(w%e(meval))
      read(w%unit) w%e
      if (ibase - 1 + ne > meval) call fatal('fewer wavefunctions in the file than requested')
      do i = 1, ibase - 1
        read(w%unit)
      end do
      call jtran_matrix(w, jay_ipar, jstart, jj0, nang)
      if (nang * nrad /= w%nbass(1)) call fatal('angular functions of the IWAVE file inconsistent')
      blk(1)%nin = w%nidvr
      allocate(blk(1)%vec(w%nidvr * nrad, ne))
      do i = 1, ne
        read(w%unit) blk(1)%vec(:, i)
      end do
    else
      read(w%unit) meval
      allocate(w%e(meval))
      read(w%unit) w%e
      if (ibase - 1 + ne > meval) call fatal('fewer vectors in the file than requested')
      do k = 1, w%jk
        nb = w%nbass(k)
        if (nb == 0) then
          read(w%unit)
          cycle
        end if
        if (mod(nb, nrad) /= 0) call fatal('block size is not a multiple of the radial grid')
        blk(k)%nin = nb / nrad
        allocate(blk(k)%vec(nb, ne))
        read(w%unit) (dum, i = 1, nb * (ibase - 1)), blk(k)%vec
      end do
    end if
  end subroutine read_body

  ! first l, offset and count of the FBR functions of a J = 0 IWAVE file
  ! (JTRAN P:1578-1600)
  subroutine jtran_matrix(w, jay_ipar, jstart, jj0, nang)
    type(wavefile), intent(in) :: w
    integer, intent(in) :: jay_ipar
    integer, intent(out) :: jstart, jj0, nang
    integer :: par
    jstart = w%kz
    jj0 = -jdia
    par = jay_ipar
    if (zembed) par = ipar1
    if (idia == 2 .and. mod(jstart, 2) /= par) then
      jj0 = jj0 + 1
      jstart = jstart + 1
    end if
    nang = (w%maxleg + w%kz - jstart) / jdia + 1
  end subroutine

  ! the values of a block's input functions at the integration points
  ! (DSRD P:1511-1548): normalised associated Legendre functions of order
  ! kz (every other one from jfirst for IDIA = 2), for a J = 0 IWAVE file
  ! multiplied by JTRAN's rows of pleg*w
  subroutine angular_values(w, b, kz, jay_ipar, blk)
    type(wavefile), intent(in) :: w
    integer, intent(in) :: b, kz, jay_ipar
    type(blockdata), intent(inout) :: blk
    double precision, allocatable :: pl(:, :), sel(:, :), psel(:, :)
    integer :: nang, jfirst, ll, jstart, jj0, j, a, m
    m = kz
    if (w%jk <= 1) then
      call jtran_matrix(w, jay_ipar, jstart, jj0, nang)
      m = w%kz                       ! the order DSRD reads from the file (P:1489)
    else
      nang = w%nbass(b) / nrad
    end if
    allocate(pl(ipot, 0:nang * jdia), sel(ipot, nang))
    call asleg(pl, nang * jdia - 1, xd, ipot, m)
    if (idia == 2) then
      if (zembed) then
        jfirst = mod(kz + ipar1, 2)
      else
        jfirst = mod(kz + jay_ipar, 2)
      end if
      do ll = 0, nang - 1
        sel(:, ll + 1) = pl(:, jfirst + 2 * ll)
      end do
    else
      sel = pl(:, 0:nang - 1)
    end if
    if (w%jk <= 1) then
      allocate(psel(nang, w%nidvr))
      do a = 1, w%nidvr
        do j = 1, nang
          psel(j, a) = w%pleg(jj0 + j * jdia, a)
        end do
      end do
      allocate(blk%ang(ipot, w%nidvr))
      blk%ang = matmul(sel, psel)
    else
      allocate(blk%ang(ipot, nang))
      blk%ang = sel
    end if
  end subroutine angular_values

  ! ASLEG (P:1234-1308): normalised P_{m+ll}^m(x_i), ll = 0 .. lmax, with
  ! the (-1)^m phase of the recurrence
  subroutine asleg(pleg, lmax, x, n, m)
    integer, intent(in) :: lmax, n, m
    double precision, intent(out) :: pleg(n, 0:lmax + 1)
    double precision, intent(in) :: x(n)
    double precision :: pmm, fact, somx2, pmmp1, pll, pnorm
    integer :: i, j, l, ll, jj
    do i = 1, n
      if (abs(x(i)) > x1) call fatal('improper argument in subroutine asleg')
      pmm = x1
      fact = x1
      somx2 = sqrt((x1 - x(i)) * (x1 + x(i)))
      do j = 1, m
        pmm = -pmm * fact * somx2
        fact = fact + x2
      end do
      pleg(i, 0) = pmm
      pmmp1 = x(i) * dble(m + m + 1) * pmm
      pleg(i, 1) = pmmp1
      ll = 1
      do l = 2 + m, lmax + m
        pll = (x(i) * dble(l + l - 1) * pmmp1 - dble(l + m - 1) * pmm) / dble(l - m)
        pmm = pmmp1
        pmmp1 = pll
        ll = ll + 1
        pleg(i, ll) = pll
      end do
    end do
    jj = -1
    do j = m, lmax + m
      fact = x1
      do i = j - m + 1, j + m
        fact = fact * dble(i)
      end do
      jj = jj + 1
      pnorm = sqrt(dble(j + j + 1) / (fact + fact))
      pleg(:, jj) = pleg(:, jj) * pnorm
    end do
  end subroutine asleg

  ! npot-point Gauss-Legendre quadrature as LAGPT and JACOBI form it
  ! (P:1166-1204, 1311-1409; Stroud and Secrest's start values for the
  ! zeros, Newton steps, the weight from the derivative of the last step's
  ! starting point), then the mirror points
  subroutine gauss_legendre(n, x, a)
    integer, intent(in) :: n
    double precision, intent(out) :: x(n), a(n)
    double precision :: b(n), c(n), fn, cc, tsa, csa, xt, xi, r1, r2, r3, dpn, pn1
    integer :: i, j
    b = x0
    c = x0
    do i = 2, n
      xi = dble(i)
      c(i) = 4.0d0 * (xi - x1) * (xi - x1) * (xi - x1) * (xi - x1) / &
             ((x2 * xi - x1) * (x2 * xi - x2) * (x2 * xi - x2) * (x2 * xi - 3.0d0))
    end do
    fn = dble(n)
    cc = x2
    tsa = cc / x2
    do i = 2, n
      cc = cc * c(i)
    end do
    csa = x0
    xt = x0
    do i = 1, n / 2
      if (i == 1) then
        r1 = 2.78d0 / (4.0d0 + fn * fn)
        xt = x1 - r1
      else if (i == 2) then
        r1 = 4.1d0
        r2 = x1 + 0.06d0 * (fn - 8.0d0) / fn
        xt = xt - r1 * r2 * (x1 - xt)
      else if (i == 3) then
        r1 = 1.67d0
        r2 = x1 + 0.22d0 * (fn - 8.0d0) / fn
        r3 = x1
        xt = xt - r1 * r2 * r3 * (x(1) - xt)
      else
        xt = 3.0d0 * x(i - 1) - 3.0d0 * x(i - 2) + x(i - 3)
      end if
      call newton(xt, n, b, c, dpn, pn1)
      x(i) = xt
      a(i) = cc / (dpn * pn1)
      csa = csa + a(i)
    end do
    write(6, 1000) n, 0, (x(i), a(i), i = 1, n / 2)
1000 format(//,i8,' point gauss-associated legendre integration',' with k =',i2, &
            //,5x,'integration points',11x,'weights',/,(f23.15,d25.12))
    write(6, 1010) csa, tsa
1010 format(/,5x,'computed sum of weights',d26.15,/,5x,'exact    sum of weights',d26.15//)
    if (abs((csa - tsa) / tsa) > 1.0d-8) call fatal('gauss-legendre weights in error: adjust algorithm')
    do i = 1, n / 2
      j = n + 1 - i
      x(j) = -x(i)
      a(j) = a(i)
    end do
  end subroutine gauss_legendre

  subroutine newton(x, n, b, c, dpn, pn1)
    double precision, intent(inout) :: x
    integer, intent(in) :: n
    double precision, intent(in) :: b(n), c(n)
    double precision, intent(out) :: dpn, pn1
    double precision :: p, dp, p1, dp1, q, dq, d
    integer :: it, j
    it = 0
    do
      it = it + 1
      p1 = x1
      p = x
      dp1 = x0
      dp = x1
      do j = 2, n
        q = (x - b(j)) * p - c(j) * p1
        dq = (x - b(j)) * dp + p - c(j) * dp1
        p1 = p
        p = q
        dp1 = dp
        dp = dq
      end do
      pn1 = p1
      d = p / dp
      x = x - d
      if (.not. (abs(d) > 1.0d-12 .and. it < 10)) exit
    end do
    dpn = dp
  end subroutine newton

  ! THREEJ (P:1778-1822) from the binomial table of SETFAC (P:1118-1138)
  double precision function threej(j1, j2, j3, m1, m2, m3, binom, nbin)
    integer, intent(in) :: j1, j2, j3, m1, m2, m3, nbin
    double precision, intent(in) :: binom(nbin, nbin)
    integer :: i1, i2, i3, k1, k2, k3, l1, l2, l3, n1, n2, n3, imin, imax, i
    double precision :: sgn
    threej = x0
    if (m1 + m2 + m3 /= 0) return
    i1 = -j1 + j2 + j3 + 1
    i2 = j1 - j2 + j3 + 1
    i3 = j1 + j2 - j3 + 1
    k1 = j1 + m1 + 1
    k2 = j2 + m2 + 1
    k3 = j3 + m3 + 1
    l1 = j1 - m1 + 1
    l2 = j2 - m2 + 1
    l3 = j3 - m3 + 1
    if (min(i1, i2, i3, k1, k2, k3, l1, l2, l3) <= 0) return
    n1 = -j1 - m2 + j3
    n2 = m1 - j2 + j3
    n3 = j1 - j2 + m3
    imin = max(-n1, -n2, 0) + 1
    imax = min(l1, k2, i3)
    if (imin > imax) return
    sgn = x1
    do i = imin, imax
      sgn = -sgn
      threej = threej + sgn * binom(i1, n1 + i) * binom(i2, n2 + i) * binom(i3, i)
    end do
    threej = threej * sqrt(binom(j2 + j2 + 1, i3) * binom(j1 + j1 + 1, i2) &
             / (binom(j1 + j2 + j3 + 2, i3) * dble(j3 + j3 + 1) &
             * binom(j1 + j1 + 1, l1) * binom(j2 + j2 + 1, l2) * binom(j3 + j3 + 1, l3)))
    if (mod(n3 + imin, 2) /= 0) threej = -threej
  end function threej

  ! the k -> k' blocks of DMAIN in the reference's order (P:877-913 for
  ! nu = 0, P:933-1049 for nu = +-1) with TRANS's factor threej * xfac *
  ! (-1)^kk1 (P:1699-1700); factor 0 marks a block with an empty side
  subroutine schedule()
    integer :: ks1, k1, k2, kk1, kk2, i, j, n
    double precision :: xfac
    nbin = jrot1 + jrot2 + 3
    allocate(binom(nbin, nbin))
    binom = x0
    binom(1, 1) = x1
    binom(2, 1) = x1
    binom(2, 2) = x1
    do i = 3, nbin
      binom(i, 1) = x1
      binom(i, i) = x1
      do j = 2, i - 1
        binom(i, j) = binom(i - 1, j - 1) + binom(i - 1, j)
      end do
    end do
    n = 3 * (jk1 + 1)
    allocate(pb1(n), pb2(n), pcomp(n), pfac(n))
    npair = 0
    ks1 = 1
    if (kmin1 > kmin2) ks1 = 2
    do k1 = ks1, jk1                          ! nu = 0
      k2 = k1 - kmin1 + kmin2
      if (k2 > jk2) cycle
      call add(k1, k2, 0, x1, 0)
    end do
    do k1 = 1, jk1                            ! nu = +1, -1
      kk1 = k1 - kmin1
      if (ket%nbass(k1) == 0) cycle
      if (jk2 > 1) then
        kk2 = kk1 + 1
        k2 = kk2 + kmin2
        if (k2 <= jk2) then
          xfac = -x1 / sqrt(x2)
          if (kk1 == 0) xfac = -x1
          if (.not. zembed .and. idia < 0) xfac = -xfac
          call add(k1, k2, 1, xfac, 1)
        end if
      end if
      kk2 = kk1 - 1
      k2 = kk2 + kmin2
      if (k2 >= 1) then
        xfac = x1 / sqrt(x2)
        if (kk2 == 0) xfac = x1
        if (.not. zembed .and. idia < 0) xfac = -xfac
        call add(k1, k2, -1, xfac, 1)
      end if
    end do
  end subroutine schedule

  subroutine add(k1, k2, nu, xfac, comp)
    integer, intent(in) :: k1, k2, nu, comp
    double precision, intent(in) :: xfac
    double precision :: f
    integer :: kk1, kk2
    kk1 = k1 - kmin1
    kk2 = k2 - kmin2
    f = threej(jk1 - kmin1, 1, jk2 - kmin2, kk1, nu, -kk2, binom, nbin) * xfac
    if (mod(kk1, 2) /= 0) f = -f
    if (ket%nbass(k1) == 0 .or. bra%nbass(k2) == 0) f = x0
    npair = npair + 1
    pb1(npair) = k1
    pb2(npair) = k2
    pcomp(npair) = comp
    pfac(npair) = f
  end subroutine add

  ! LAGPT (P:1206-1230): dipol(j + ipot (i1 - 1 + npnt1 (i2 - 1)), nu + 1)
  ! = wtd_j DIPD(r1_i1, r2_i2, x_j, nu), doubled for IDIA = 2
  subroutine lagpt(nu)
    integer, intent(in) :: nu
    integer :: i1, i2, j, ii
    double precision :: d
    ii = 0
    do i2 = 1, bra%npnt2
      do i1 = 1, bra%npnt1
        do j = 1, ipot
          ii = ii + 1
          call dipd(d, bra%r1(i1), bra%r2(i2), xd(j), nu)
          if (jdia == 2) d = x2 * d
          dipol(ii, nu + 1) = wtd(j) * d
        end do
      end do
    end do
  end subroutine lagpt

  subroutine dmain()
    type(dvr_dip_dims_t) :: dims
    type(dvr_dip_block_t), allocatable, target :: ck(:), cb(:)
    type(dvr_dip_pair_t), allocatable :: cp(:)
    integer :: q, iblock, kblock, nrun, rc, k, first
    logical, allocatable :: usek(:), useb(:)
    call read_body(ket, ne1, ibase1, ipar1, kblk)
    call read_body(bra, ne2, ibase2, ipar2, bblk)
    write(6, '(///,5x,a,/)') 'First 10 energies for the ket in a.u.'
    if (.not. zpmin) write(6, '(5d24.12,/)') ket%e(ibase1:ibase1 - 1 + min(10, ne1))
    write(6, '(///,5x,a,/)') 'First 10 energies for the bra in a.u.'
    if (.not. zpmin) write(6, '(5d24.12,/)') bra%e(ibase2:ibase2 - 1 + min(10, ne2))
    allocate(tz(ne1, ne2), tx(ne1, ne2))
    tz = x0
    tx = x0
    iblock = 0
    if (ires > 0) then
      open(unit=iscr, form='unformatted', status='old', action='read')
      read(iscr) iblock
      read(iscr) tz
      read(iscr) tx
      close(iscr)
    end if
    kblock = iblock + nblock
    allocate(xd(npot), wtd(npot), dipol(nrad * ipot, 2))
    write(6, '(/,5x,a)') 'lagpt called'
    call gauss_legendre(npot, xd, wtd)
    call lagpt(0)
    call lagpt(1)
    call schedule()
    ! the blocks this run computes: iblock+1 .. min(kblock, mblock) of the list
    first = iblock + 1
    nrun = max(0, min(kblock, npair) - iblock)
    allocate(usek(jk1), useb(jk2), ck(jk1), cb(jk2), cp(max(1, nrun)))
    usek = .false.
    useb = .false.
    do q = 1, nrun
      k = first + q - 1
      cp(q)%b1 = pb1(k) - 1
      cp(q)%b2 = pb2(k) - 1
      cp(q)%comp = pcomp(k)
      cp(q)%factor = pfac(k)
      if (pfac(k) /= x0) then
        usek(pb1(k)) = .true.
        useb(pb2(k)) = .true.
      end if
    end do
    do k = 1, jk1
      ck(k)%nin = 0
      ck(k)%vec = c_null_ptr
      ck(k)%ang = c_null_ptr
      if (.not. usek(k)) cycle
      call angular_values(ket, k, k - kmin1, ipar1, kblk(k))
      if (ket%jk <= 1) then
        ck(k)%nin = ket%nidvr
      else
        ck(k)%nin = kblk(k)%nin
      end if
      ck(k)%vec = c_loc(kblk(k)%vec)
      ck(k)%ang = c_loc(kblk(k)%ang)
    end do
    do k = 1, jk2
      cb(k)%nin = 0
      cb(k)%vec = c_null_ptr
      cb(k)%ang = c_null_ptr
      if (.not. useb(k)) cycle
      call angular_values(bra, k, k - kmin2, ipar2, bblk(k))
      if (bra%jk <= 1) then
        cb(k)%nin = bra%nidvr
      else
        cb(k)%nin = bblk(k)%nin
      end if
      cb(k)%vec = c_loc(bblk(k)%vec)
      cb(k)%ang = c_loc(bblk(k)%ang)
    end do
    dims%nrad = nrad
    dims%ipot = ipot
    dims%ne1 = ne1
    dims%ne2 = ne2
    dims%nblk1 = jk1
    dims%nblk2 = jk2
    dims%npair = nrun
    rc = dvr_dipole3(dims, ck, cb, dipol, cp, tz, tx, c_null_ptr)
    if (rc /= DVR_OK) then
      write(6, '(/5x,a,i3)') 'dvr_dipole3 failed with code', rc
      stop 1
    end if
    do q = 1, nrun
      k = first + q - 1
      if (pfac(k) == x0) then
        write(6, 2020) k, pb1(k) - kmin1, pb2(k) - kmin2
      else
        write(6, 2000) k, pb1(k) - kmin1, pb2(k) - kmin2
      end if
    end do
2000 format(/5x,'Block',i4,' k1 =',i3,' to k2 =',i3,' completed')
2020 format(/5x,'Block',i4,' k1 =',i3,' to k2 =',i3,' skipped')
    iblock = iblock + nrun
    open(unit=iscr, file=unit_file(iscr), form='unformatted', status='replace', action='write')
    write(iscr) iblock
    write(iscr) tz
    write(iscr) tx
    close(iscr)
    if (iblock < mblock) write(6, '(//i7,a)') iblock, ' blocks calculated. dipole3 shutting down'
    call spect()
  end subroutine dmain

  ! SPECT (P:1825-1953) and OUTPT (P:1956-2002)
  subroutine spect()
    double precision :: ezero, xf, xe1, dd, dd3, tzd, txd, t, sxd, a
    double precision, allocatable :: xe2(:), sint(:, :)
    integer :: j1, j2, ie1, ie2, ios
    j1 = jk1 - kmin1
    j2 = jk2 - kmin2
    write(6, '(///)')
    write(6, 201)
201 format(//,5x,'*************************************************' &
           //,5x,'print out of dipole transition moments and s(f-1)' &
           //,9x,'frequencies in wavenumbers', &
            /,9x,'transition moments in debye (2.54174a.u.)', &
            /,9x,'s(f-i) in debye**2', &
            /,9x,'einstein a-coefficient in sec-1',//, &
           5x,'*************************************************')
    write(6, '(///)')
    write(6, '(5x,9a8)') title
    write(6, '(///)')
    write(6, 203) j1, kmin1, j2, kmin2, idia, ipar1, ipar2
203 format(5x,'j1=',i4,'  kmin1=',i4,'  j2=',i4,'  kmin2=',i4, &
           '  idia=',i4,'  ipar1=',i4,'  ipar2=',i4,//)
    ezero = x0
    read(5, '(f20.0)', iostat=ios) ezero
    if (ios /= 0) then
      ezero = x0
    else
      write(6, '(5x,a,e16.8,a)') 'ground zero =', ezero, ' cm-1'
    end if
    write(6, '(///)')
    write(6, 205)
205 format(/,' ie1 ie2   ket energy   bra energy    frequency  z trans', &
           'ition    x transition       dipole       s(f-i)      a-coefficient',/)
    xf = sqrt(dble((2 * j1 + 1) * (2 * j2 + 1)))
    allocate(xe2(ne2), sint(ne1, ne2))
    do ie2 = 1, ne2
      xe2(ie2) = bra%e(ibase2 - 1 + ie2) * autocm - ezero
    end do
    a = x0
    do ie1 = 1, ne1
      xe1 = ket%e(ibase1 - 1 + ie1) * autocm - ezero
      do ie2 = 1, ne2
        dd = xe2(ie2) - xe1
        dd3 = abs(dd * dd * dd)
        sint(ie1, ie2) = (tz(ie1, ie2) + tx(ie1, ie2))**2 * xf * xf
        tzd = tz(ie1, ie2) * autode * xf
        txd = tx(ie1, ie2) * autode * xf
        if (.not. zbisc .and. zembed) txd = -txd
        t = abs(tzd + txd)
        sxd = t * t
        if (dd > x0) a = sxd * dd3 * detosec / dble(2 * j2 + 1)
        if (dd < x0) a = sxd * dd3 * detosec / dble(2 * j1 + 1)
        if (zpmin .and. max(ie1, ie2) > 10) cycle
        write(6, 206) ie1, ie2, xe1, xe2(ie2), dd, tzd, txd, t, sxd, a
206     format(2(i4),3(3x,f14.6),5(2x,es15.8))
      end do
    end do
    if (ztra) then
      if (zstart) then
        open(unit=itra, file=unit_file(itra), form='unformatted', status='replace', &
             action='write')
      else
        open(unit=itra, file=unit_file(itra), form='unformatted', position='append', &
             action='write')
      end if
      write(itra) j1, j2, kmin1, kmin2, ne1, ne2, idia, ipar1, ipar2, ezero, zembed, &
                  ibase1, ibase2
      write(itra) ket%e(ibase1:ibase1 - 1 + ne1)
      write(itra) bra%e(ibase2:ibase2 - 1 + ne2)
      do ie2 = 1, ne2
        write(itra) sint(:, ie2)
      end do
      close(itra)
    end if
    write(6, '(/,5x,a)') 'program ended'
  end subroutine spect

end program dipole3_amd
