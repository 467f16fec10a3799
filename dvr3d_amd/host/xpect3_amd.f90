! Drop-in XPECT3 on the MI355X: expectation values of user properties over
! the vib-rot wavefunctions of one run (source/xpect3.f90, cited as X:<line>).
! Same namelist PRT and cards on stdin, the same ket file (the IWAVE file of
! a J = 0 DVR3DRJZ run, jhmain or nfmain, or the KVEC file of ROTLEV3), the
! same stdout table and ITRA/ITRA0 files; the evaluation of the wavefunctions
! on the grid and the sums over it run on the GPU through dvr_xpect3
! (libdvr3d_amd).
!
! The host does what the reference's host does: INSIZE (X:76-295), GENIND
! (X:320-360), the grids and energies of X3MAIN (X:380-421), the property
! routine on the grid (X:438-505 for J = 0, LAGPT X:562-657 for J > 0, with
! the sign-scan Gauss-Legendre rule of JACOBI X:659-758), REST's angular
! values (X:848-925) and OUTPT2 (X:1086-1280).  The properties are the
! user's SUBROUTINE PROP(P,R1,R2,XCOS,NPROPIN,NPRT), linked as for the
! reference, with COMMON /MASS/ and /POT/ IPROP filled as the reference
! fills them.  Not supported (refused with a message): J = 0 files with
! IDIA > 0 (the reference reads their energy-count record as the angular
! grid) and non-Coriolis J > 0 files.  With zform = F and zfit = F the
! energies are written as one unformatted record (the reference writes them
! list-directed to the unformatted unit, which its runtime refuses).
program xpect3_amd
  use iso_c_binding
  use dvr3d_amd_api
  implicit none

  ! namelist PRT (X:18-19) with the block data defaults (X:68-71)
  logical :: zmors1, zncor, zprint, ztra, zstart, zmors2, zfit, zform
  integer :: iket, itra, itra0, ilev
  namelist /prt/ zprint, ztra, zstart, zfit, zform, iket, itra, itra0, ilev
  double precision :: xmass(3), g1, g2
  logical :: zembed, zbisc
  common /mass/ xmass, g1, g2, zembed, zbisc
  integer, parameter :: mxprop = 1000
  integer :: iprop
  common /pot/ iprop(mxprop)

  double precision, parameter :: x0 = 0.0d0, xp5 = 0.5d0, autocm = 2.19474624d+05
  character(len=8) :: title(9)
  integer :: idia, ipar, lmax, npnt1, npnt2, jrot, kmin1, neval, jk1, mbass, nrade
  integer :: lpot, npot, npropin, nprt, nv1, nn2, meval
  integer, allocatable :: nbass1(:), lmin1(:), lbass1(:), iv(:)
  double precision :: re1, diss1, we1, re2, diss2, we2
  double precision, allocatable :: r1(:), r2(:), xalf(:), eh(:), props(:, :)
  double precision, allocatable :: xd(:), wtd(:), cj(:)
  integer :: i

  type blockvals
    double precision, allocatable :: vec(:, :), ang(:, :)
  end type

  zncor = .false.; zprint = .false.; zfit = .false.; ztra = .true.; zstart = .false.
  zform = .true.; iket = 11; itra = 12; itra0 = 28; ilev = 14
  zmors1 = .true.; zmors2 = .true.

  write(6, 200)
200 format(//5x,'program xpect3 (MI355X / gfx950 build of the expectation-value step):',/)
  read(5, prt)
  read(5, '(9a8)') title
  write(6, '(/5x,9a8)') title
  call insize()
  call x3main()

contains

  subroutine fatal(msg)
    character(len=*), intent(in) :: msg
    write(6, '(//,5x,a,/)') msg
    stop 1
  end subroutine

  character(len=16) function unit_file(iu)
    integer, intent(in) :: iu
    write(unit_file, '(a,i0)') 'fort.', iu
  end function

  ! INSIZE (X:76-295)
  subroutine insize()
    integer :: ios, nmax11, nmax12, mbass0, jt
    if (zfit) write(6, '(/5x,a)') 'Property analysis for potential energy fitting'
    if (.not. zfit) write(6, '(/5x,a)') 'Property analysis to obtain expectation values'
    write(6, '(//5x,a,i4)') 'input ket wavefunctions read from stream iket  =', iket
    if (ztra) then
      write(6, '(5x,a,i4)') 'output transition data written to stream itra  =', itra
      if (zform) write(6, '(5x,a)') 'file written formatted'
      if (.not. zform) write(6, '(5x,a)') 'file written unformatted'
      if (.not. zfit) then
        if (zstart) write(6, '(5x,a)') 'file written from the start'
        if (.not. zstart) write(6, '(5x,a)') 'data appended to the end of file'
      end if
    end if
    open(unit=iket, form='unformatted', status='old', action='read', iostat=ios)
    if (ios /= 0) call fatal('cannot open the wavefunction file')
    read(iket) idia, ipar, lmax, nmax11, nmax12, jrot, kmin1, neval
    read(iket) zembed, zmors1, zmors2, xmass, g1, g2, zncor
    read(iket) re1, diss1, we1, re2, diss2, we2
    if (idia > -2) then
      zbisc = .false.
    else
      zbisc = .true.
      zembed = .true.
    end if
    if (zfit .and. max(jrot, ipar) > 0) write(6, '(5x,a,i4)') &
      'Ground state data read from       stream itra0 =', itra0
    write(6, '(5x,a,i4)') 'Ground state energy read from     stream ilev   ', ilev
    npnt1 = nmax11
    npnt2 = nmax12
    read(5, '(5i5)') lpot, npropin, nprt, nv1
    if (nv1 > neval) call fatal('NV1 > NEVAL! (NV1 MUST BE <= NEVAL)')
    if (nv1 == 0) nv1 = neval
    if (npropin == 0) call fatal('NO PROPERTIES TO BE CONSIDERED! (NPROPIN = 0)')
    if (npropin > mxprop) call fatal('more properties requested than mxprop')
    if (nprt < npropin) nprt = npropin
    if (lpot == 0) lpot = lmax + jrot + mod(lmax + jrot, 2)
    if (jrot == 0) lpot = lmax
    iprop = 0
    read(5, '(20i5)') (iprop(i), i = 1, npropin)
    if (iprop(1) <= 0) then
      do i = 1, npropin
        iprop(i) = i
      end do
    end if
    write(6, 201) npropin, nprt, lpot, (iprop(i), i = 1, npropin)
201 format(/,5x,'Number of properties to be considered   =',i5, &
           /,5x,'Number of properties computed in PROP   =',i5, &
           /,5x,'Number of angular DVR points            =',i5, &
           /,5x,'properties to be considered are numbered:',/,5x,(25i5))
    if (jrot == 0) kmin1 = 1
    if (zncor) then
      if (jrot /= 0) call fatal('non-Coriolis J > 0 wavefunctions are not supported')
      jk1 = 1
    else
      jk1 = jrot + kmin1
    end if
    if (idia > -2) then
      nrade = npnt1 * npnt2
      mbass = nrade * jk1 * lmax
    else
      nrade = npnt1 * (npnt1 + 1) / 2
      jt = jk1 / 2
      mbass = (nrade + npnt1 * (npnt1 - 1) / 2) * jt
      if (2 * jt /= jk1) then
        if (ipar == 0) mbass = mbass + nrade
        if (ipar /= 0) mbass = mbass + npnt1 * (npnt1 - 1) / 2
      end if
      mbass = mbass * lmax
    end if
    write(6, 300) mbass, nmax11, nmax12, lmax, jrot, ipar, kmin1, neval, nv1
300 format(5x,'total number of basis functions   = ',i5,/ &
           5x,'number of r1 radial dvr points    = ',i5,/ &
           5x,'number of r2 radial dvr points    = ',i5,/ &
           5x,'number of angular functions       = ',i5,/ &
           5x,'total angular momentum            = ',i5,/ &
           5x,'ipar, vibrational parity          = ',i5,/ &
           5x,'kmin= 1-p, where j + p is parity, = ',i5,/ &
           5x,'total vib-rot functions available = ',i5,/ &
           5x,'vib-rot functions taken (0 = all) = ',i5,/)
    npot = lpot
    if (jrot > 0) then
      jt = lpot
      if (idia == 2 .and. mod(jt, 2) /= ipar) jt = jt + 1
      npot = ((jt + 2) / 2) * 2
      nn2 = npot / 2
    end if
    ! GENIND (X:320-360)
    allocate(nbass1(jk1), lmin1(jk1), lbass1(jk1))
    read(iket, iostat=ios) mbass0, lmin1, lbass1, nbass1
    if (ios /= 0 .or. mbass0 > mbass) then
      write(6, '(//,5x,a,/,5x,a,i3,a,i6,a,i6,a)') 'basis function dimensions in error', &
        'unit =', iket, ' mbass =', mbass, ' expected, =', mbass0, ' found'
      stop 1
    end if
  end subroutine insize

  ! X3MAIN (X:363-556)
  subroutine x3main()
    type(dvr_xp_dims_t) :: dims
    type(dvr_xp_block_t), allocatable :: cb(:)
    type(blockvals), allocatable, target :: bv(:)
    double precision, allocatable, target :: d0(:, :)
    integer :: k, rc, kz, nread, iang, ibass, i2, i1, j, ii, i0
    double precision, allocatable :: prin(:), trin(:)
    allocate(r1(npnt1), r2(npnt2), prin(npropin), trin(npropin))
    read(iket) r1
    if (idia > -2) then
      read(iket) r2
    else
      r2 = r1
    end if
    if (jk1 <= 1) then
      if (idia >= 0) call fatal('J = 0 wavefunctions of IDIA > 0 cannot be read (XPECT3 reads '// &
                                'their energy-count record as the angular grid)')
      allocate(xalf(lmax))
      read(iket) xalf
      read(iket)                                 ! kz, maxleg, nidvr, lincr
      read(iket)                                 ! pleg * w
      if (idia == -2) then
        read(iket) iang, ibass
        allocate(iv(lmax))
        read(iket) iv
      else
        ibass = mbass
      end if
    end if
    read(iket) meval
    allocate(eh(meval))
    read(iket) eh
    if (nv1 > meval) call fatal('fewer levels in the file than NV1')
    allocate(props(nv1, npropin))
    props = x0
    dims%ne = nv1
    dims%nprop = npropin
    if (jrot == 0) then
      ! J = 0: the DVR amplitudes are the values at the grid points; the
      ! properties on the same grid (X:438-505)
      allocate(bv(1), cb(1))
      allocate(bv(1)%vec(ibass, nv1))
      do i = 1, nv1
        read(iket) bv(1)%vec(:, i)
      end do
      if (idia == -2) then
        allocate(d0((npnt2 * (npnt2 + 1 - 2 * ipar)) / 2 * count(iv /= 0), npropin))
      else
        allocate(d0(nrade * lmax, npropin))
      end if
      ii = 0
      do i2 = 1, npnt2
        i0 = npnt1
        if (idia == -2) i0 = i2 - ipar
        do i1 = 1, i0
          do j = 1, lmax
            if (idia == -2) then
              if (iv(j) == 0) cycle
            end if
            ii = ii + 1
            call prop(prin, r1(i1), r2(i2), xalf(j), npropin, nprt)
            if (idia == -2) then
              call prop(trin, r2(i2), r1(i1), xalf(j), npropin, nprt)
              d0(ii, :) = xp5 * (prin + trin)
            else
              d0(ii, :) = prin
            end if
          end do
        end do
      end do
      if (ii > ibass) call fatal('grid larger than the wavefunction records')
      dims%npts = ii
      dims%npot = 0
      dims%nblk = 1
      cb(1)%nin = 0
      cb(1)%ldv = ibass
      cb(1)%vec = c_loc(bv(1)%vec)
      cb(1)%ang = c_null_ptr
    else
      ! J > 0: the properties at the quadrature points (LAGPT) and every k
      ! block's coefficients with REST's angular values
      allocate(xd(npot), wtd(npot), d0(nrade * npot, npropin))
      call lagpt(d0)
      allocate(bv(jk1), cb(jk1))
      dims%npts = nrade * npot
      dims%npot = npot
      dims%nblk = jk1
      do k = 1, jk1
        cb(k)%nin = 0
        cb(k)%ldv = 0
        cb(k)%vec = c_null_ptr
        cb(k)%ang = c_null_ptr
        if (nbass1(k) == 0) then
          read(iket)
          cycle
        end if
        nread = nbass1(k)
        if (nread /= nrade * lmax) call fatal('k block size is not lmax radial points times')
        allocate(bv(k)%vec(nread, nv1))
        read(iket) bv(k)%vec
        kz = k - kmin1
        call rest(bv(k), kz)
        cb(k)%nin = lmax
        cb(k)%ldv = nread
        cb(k)%vec = c_loc(bv(k)%vec)
        cb(k)%ang = c_loc(bv(k)%ang)
      end do
    end if
    rc = dvr_xpect3(dims, cb, d0, props)
    if (rc /= DVR_OK) then
      write(6, '(/5x,a,i3)') 'dvr_xpect3 failed with code', rc
      stop 1
    end if
    write(6, '(/,5x,a)') 'trans called'
    call outpt2()
  end subroutine x3main

  ! LAGPT (X:562-657): the sign-scan Gauss-Legendre rule, mirrored and
  ! reflected, and the properties at (r1, r2, x), averaged over x -> -x for
  ! IDIA = 2
  subroutine lagpt(d0)
    double precision, intent(out) :: d0(:, :)
    double precision :: cc, tsa, csa, xxx, pm1, p, dp, pn1, d, t
    double precision :: prin(npropin), trin(npropin)
    double precision :: xt(npot)
    real :: step
    integer :: i, j, ii, nf, it, i1, i2, jj
    allocate(cj(npot))
    cj = x0
    do i = 2, npot
      t = dble(i)
      cj(i) = 4.0d0 * (t - 1.0d0) * (t - 1.0d0) * (t - 1.0d0) * (t - 1.0d0) / &
             ((2.0d0 * t - 1.0d0) * (2.0d0 * t - 2.0d0) * (2.0d0 * t - 2.0d0) * (2.0d0 * t - 3.0d0))
    end do
    cc = 2.0d0
    tsa = cc
    do i = 2, npot
      cc = cc * cj(i)
    end do
    step = 0.0001
    nf = 0
    pm1 = 1.0d0
    xxx = 1.0d0
    do
      call recur(xxx, p, dp, pn1)
      if (pm1 * p < x0) then
        pm1 = -pm1
        nf = nf + 1
        xt(nf) = xxx
      end if
      if (nf == nn2) exit
      xxx = xxx - step
      if (.not. (xxx > -0.0002)) call fatal('Incorrect number of zeros found in JACOBI')
    end do
    csa = x0
    do i = 1, nn2
      t = xt(i)
      it = 0
      do
        it = it + 1
        call recur(t, p, dp, pn1)
        d = p / dp
        t = t - d
        if (.not. (abs(d) > 1.0d-12 .and. it < 10)) exit
      end do
      xd(i) = t
      wtd(i) = cc / (dp * pn1)
      csa = csa + wtd(i) + wtd(i)
    end do
    if (2 * nn2 /= npot) csa = csa - wtd(nn2)
    write(6, 1000) npot, 0, (xd(i), wtd(i), i = 1, nn2)
1000 format(//,i8,' point gauss-associated legendre integration',' with k =',i2, &
            //,5x,'integration points',11x,'weights',/,(f23.15,d25.12))
    write(6, 1010) csa, tsa
1010 format(/,5x,'computed sum of weights',d26.15,/,5x,'exact    sum of weights',d26.15//)
    if (abs((csa - tsa) / tsa) > 1.0d-8) call fatal('gauss-legendre weights in error: adjust algorithm')
    do i = 1, nn2
      j = i + nn2
      xd(j) = -xd(nn2 + 1 - i)
      wtd(j) = wtd(nn2 + 1 - i)
    end do
    xd = -xd
    ii = 0
    do i2 = 1, npnt2
      do i1 = 1, npnt1
        do jj = 1, npot
          ii = ii + 1
          call prop(prin, r1(i1), r2(i2), xd(jj), npropin, nprt)
          if (idia == 2) then
            call prop(trin, r1(i1), r2(i2), -xd(jj), npropin, nprt)
            d0(ii, :) = xp5 * (prin + trin)
          else
            d0(ii, :) = prin
          end if
        end do
      end do
    end do
  end subroutine lagpt

  ! the three-term recurrence of the Legendre polynomials with their
  ! derivative (RECUR X:740-758, alf = bta = 0)
  subroutine recur(x, pn, dpn, pnm1)
    double precision, intent(in) :: x
    double precision, intent(out) :: pn, dpn, pnm1
    double precision :: p1, q, dq, dp1
    integer :: jr
    p1 = 1.0d0
    pn = x
    dp1 = x0
    dpn = 1.0d0
    do jr = 2, npot
      q = x * pn - cj(jr) * p1
      dq = x * dpn + pn - cj(jr) * dp1
      p1 = pn
      pn = q
      dp1 = dpn
      dpn = dq
    end do
    pnm1 = p1
  end subroutine recur

  ! REST (X:848-925): sqrt(w_p) P_l^kz(x_p) for the block's lmax functions,
  ! l from kz (every other one from the block's lmin for IDIA = 2)
  subroutine rest(b, kz)
    type(blockvals), intent(inout) :: b
    integer, intent(in) :: kz
    double precision, allocatable :: pl(:, :)
    integer :: jdia, nang, jj, j, p
    jdia = max(1, idia)
    nang = lmax * jdia
    allocate(pl(npot, 0:nang), b%ang(npot, lmax))
    call asleg(pl, nang - 1, xd, npot, kz)
    if (idia == 2) then
      ! REST takes the first l from lmin1(kz + 1) (X:889), which is this
      ! block's for KMIN = 1
      if (kz + 1 > jk1) call fatal('IDIA = 2 block beyond the lmin table (KMIN = 0 file)')
      jj = lmin1(kz + 1) - kz
      do j = 1, lmax
        b%ang(:, j) = pl(:, jj)
        jj = jj + jdia
      end do
    else
      b%ang = pl(:, 0:lmax - 1)
    end if
    do p = 1, npot
      b%ang(p, :) = sqrt(wtd(p)) * b%ang(p, :)
    end do
  end subroutine rest

  ! normalised P_{m+ll}^m(x_i), ll = 0 .. lmax, with the (-1)^m phase of the
  ! recurrence (ASLEG X:928-996)
  subroutine asleg(pleg, lmax_, x, n, m)
    integer, intent(in) :: lmax_, n, m
    double precision, intent(out) :: pleg(n, 0:lmax_ + 1)
    double precision, intent(in) :: x(n)
    double precision :: pmm, fact, somx2, pmmp1, pll, pnorm
    integer :: i, j, l, ll, jj
    do i = 1, n
      if (abs(x(i)) > 1.0d0) call fatal('improper argument in subroutine asleg')
      pmm = 1.0d0
      fact = 1.0d0
      somx2 = sqrt((1.0d0 - x(i)) * (1.0d0 + x(i)))
      do j = 1, m
        pmm = -pmm * fact * somx2
        fact = fact + 2.0d0
      end do
      pleg(i, 0) = pmm
      pmmp1 = x(i) * dble(m + m + 1) * pmm
      pleg(i, 1) = pmmp1
      ll = 1
      do l = 2 + m, lmax_ + m
        pll = (x(i) * dble(l + l - 1) * pmmp1 - dble(l + m - 1) * pmm) / dble(l - m)
        pmm = pmmp1
        pmmp1 = pll
        ll = ll + 1
        pleg(i, ll) = pll
      end do
    end do
    jj = -1
    do j = m, lmax_ + m
      fact = 1.0d0
      do i = j - m + 1, j + m
        fact = fact * dble(i)
      end do
      jj = jj + 1
      pnorm = sqrt(dble(j + j + 1) / (fact + fact))
      pleg(:, jj) = pleg(:, jj) * pnorm
    end do
  end subroutine asleg

  ! OUTPT2 (X:1086-1280)
  subroutine outpt2()
    double precision :: ezero, ezero2, xeh, xe1
    double precision, allocatable :: xprop(:, :), pground(:), ehtmp(:)
    character(len=2) :: s1
    integer :: ie, n, ios, isym
    write(6, 203) jrot, kmin1, idia, ipar
203 format(//5x,'j1=',i4,'  kmin1=',i4,'  idia=',i4,'  ipar=',i4,//)
    ezero2 = x0
    open(unit=ilev, file=unit_file(ilev), status='old', action='read', iostat=ios)
    if (ios == 0) then
      read(ilev, *, iostat=ios)
      if (ios == 0) read(ilev, *, iostat=ios)
      if (ios == 0) read(ilev, '(d20.12)', iostat=ios) ezero2
      if (ios /= 0) ezero2 = x0
      close(ilev)
    end if
    ezero = ezero2 * autocm
    write(6, '(5x,a,d16.8)') 'ground zero in hartrees ', ezero2
    write(6, '(5x,a,d16.8,a)') 'ground zero =', ezero, ' cm-1'
    s1 = '  '
    if (ipar == 0 .and. kmin1 == 1) then
      s1 = 'a1'
    else if (ipar == 1 .and. kmin1 == 0) then
      s1 = 'b2'
    else if (ipar == 1 .and. kmin1 == 1) then
      s1 = 'a2'
    else if (ipar == 0 .and. kmin1 == 0) then
      s1 = 'b1'
    end if
    if (ztra .and. zform) then
      if (zstart .or. zfit) then
        open(unit=itra, file=unit_file(itra), form='formatted', status='replace')
      else
        open(unit=itra, file=unit_file(itra), form='formatted', position='append')
      end if
    else if (ztra) then
      if (zstart .or. zfit) then
        open(unit=itra, file=unit_file(itra), form='unformatted', status='replace')
      else
        open(unit=itra, file=unit_file(itra), form='unformatted', position='append')
      end if
    end if
    write(6, '(//,a)') ' J     ie       energy   property'
    allocate(xprop(nv1, npropin), pground(npropin), ehtmp(nv1))
    if (ipar == 0 .and. jrot == 0 .and. zfit) then
      do ie = 2, nv1
        xprop(ie, :) = props(ie, :) - props(1, :)
      end do
      open(unit=itra0, file=unit_file(itra0), form='formatted', status='replace')
      do n = 1, npropin
        write(itra0, *) props(1, n)
      end do
      close(itra0)
      do ie = 1, nv1 - 1
        xeh = eh(ie + 1) - ezero2
        xe1 = xeh * autocm
        if (ztra) write(itra, 308) jrot, s1, ie, xe1, xeh
        if (ztra) then
          do n = 1, npropin
            write(itra, *) xprop(ie + 1, n)
          end do
        end if
        write(6, 308) jrot, s1, ie, xe1, (xprop(ie + 1, n), n = 1, npropin)
      end do
    else if (zfit) then
      ! ground-state values written by the J = 0 run (X:1180-1190)
      if (zform) then
        open(unit=itra0, file=unit_file(itra0), form='formatted', status='old', iostat=ios)
        if (ios /= 0) call fatal('ground state data (itra0) missing')
        read(itra0, *) pground
      else
        open(unit=itra0, file=unit_file(itra0), form='unformatted', status='old', iostat=ios)
        if (ios /= 0) call fatal('ground state data (itra0) missing')
        read(itra0)
        read(itra0) pground
      end if
      close(itra0)
      do ie = 1, nv1
        xprop(ie, :) = props(ie, :) - pground
      end do
      do ie = 1, nv1
        xeh = eh(ie) - ezero2
        xe1 = xeh * autocm
        if (zform) then
          if (ztra) write(itra, 308) jrot, s1, ie, xe1, xeh
          do n = 1, npropin
            if (ztra .and. xprop(ie, n) /= x0) write(itra, *) xprop(ie, n)
          end do
        else
          if (ztra) write(itra) jrot, s1, ie, xe1, xeh
          do n = 1, npropin
            if (ztra .and. xprop(ie, n) /= x0) write(itra) xprop(ie, n)
          end do
        end if
        write(6, 308) jrot, s1, ie, xe1, (xprop(ie, n), n = 1, npropin)
      end do
    else
      do ie = 1, nv1
        xeh = eh(ie) - ezero2
        xe1 = xeh * autocm
        write(6, 308) jrot, s1, ie, xe1, (props(ie, n), n = 1, npropin)
      end do
      if (ztra) then
        if (idia > 0) then
          isym = 0
        else
          isym = abs(idia)
          if (ipar == 1) isym = -isym
        end if
        ehtmp = eh(1:nv1) * autocm
        if (zform) then
          write(itra, '(i4,1x,i4,1x,i4,1x,i4)') jrot, nv1, ipar, kmin1
          write(itra, *) npropin, max(idia, 0), isym, ezero
          write(itra, *) (iprop(n), n = 1, npropin)
          write(itra, *) ehtmp
          do n = 1, npropin
            write(itra, *) props(:, n)
          end do
        else
          write(itra) jrot, neval, ipar, kmin1
          write(itra) npropin, max(idia, 0), isym, ezero
          write(itra) (iprop(n), n = 1, npropin)
          write(itra) ehtmp
          do n = 1, npropin
            write(itra) props(:, n)
          end do
        end if
      end if
    end if
308 format(i2,1x,a2,i4,3x,f10.3,5(2x,d24.16)/(17x,5(2x,d13.6)))
    if (ztra) close(itra)
  end subroutine outpt2

end program xpect3_amd
