! Drop-in ROTLEV3 on the MI355X: the rotational step that reads the IWAVE
! file of a DVR3DRJZ zrot run (source/rotlev3.f90, cited as R:<line>).
! Same namelist PRT and cards on stdin, the same stdout report, ILEV, JVEC
! and KVEC files; the Hamiltonian construction and its solve run on the GPU
! through dvr_rotlev3 (libdvr3d_amd).
!
! The host does what the reference's host does: INSIZE (R:112-298), the
! energy selection of SELECT (R:301-478), the file layouts of DGROT,
! DSTORE and DSTORE1 (R:1046-1475).  Both diagonaliser choices (zdcore)
! give the lowest NEVAL levels of the same matrix, here from the GPU's
! dense solver.  Not supported (refused with a message): restart runs
! (ires /= 0), Hamiltonian-only runs (zdiag = F), DVR1DA input (z1da), and
! J = 1 with KMIN = 0, whose reference path passes an unallocated array.
program rotlev3_amd
  use iso_c_binding
  use dvr3d_amd_api
  implicit none

  ! namelist PRT and its defaults (R:97-103)
  double precision :: toler, thresh
  integer :: ilev, iwave, jscr, jvec, jvec2, kvec, kvec2, iscr, ires, irf1, irf2
  logical :: zpham, zpvec, zdcore, zvec, zpfun, zdiag, ztran, zptra
  namelist /prt/ toler, thresh, ilev, iwave, jscr, jvec, jvec2, kvec, kvec2, iscr, ires, &
                 irf1, irf2, zpham, zpvec, zdcore, zvec, zpfun, zdiag, ztran, zptra

  type kblock
    double precision, allocatable :: xalf(:), pleg(:, :), eval(:), vecs(:, :)
    integer :: kz, maxleg, idvr, lincr, meval
  end type

  double precision, parameter :: autocm = 2.19474624d+05, x0 = 0.0d0
  integer :: idia, ipar, npntt, npnt1, npnt2, jrot, kmin0, meval, nlim
  logical :: zembed, zmors1, zmors2, z1da, zquad2
  double precision :: xmass(3), g1, g2, morse(6), ezero
  double precision, allocatable :: rm2(:), r1(:), r2(:)
  type(kblock), allocatable :: blk(:)
  integer :: nvib, neval, kmin, ibass, neval2, nblk, nbass, mbass, kbass, nfile
  integer, allocatable :: mvib(:), jst(:), nang(:)
  character(len=8) :: title(9)
  integer :: i, ios, k

  toler = 0.0d0; thresh = 0.1d0; zpham = .false.; zpvec = .false.; iwave = 26
  zvec = .false.; jvec = 3; jvec2 = 2; iscr = 10; ires = 0; ztran = .false.; kvec = 8
  kvec2 = 9; zptra = .false.; jscr = 7; zpfun = .false.; ilev = 14; zdiag = .true.
  zdcore = .false.; irf1 = 21; irf2 = 22

  write(6, 1000)
1000 format(5x,'PROGRAM ROTLEV3 (MI355X / gfx950 build of the rotational step):',/)
  read(5, prt)
  call insize()
  call select_basis()
  call vrmain()

contains

  subroutine insize()
    open(unit=iwave, form='unformatted', status='old', iostat=ios)
    if (ios /= 0) then
      write(6, '(/5x,a,i4)') 'cannot open IWAVE on stream ', iwave
      stop 1
    end if
    read(iwave) idia, ipar, npntt, npnt1, npnt2, jrot, kmin0, meval, nlim
    read(iwave) zembed, zmors1, zmors2, xmass, g1, g2, z1da, zquad2
    read(iwave) morse
    allocate(rm2(nlim), r1(npnt1), r2(npnt2))
    read(iwave) rm2
    read(iwave) r1
    read(iwave) r2
    ! every k block of the file (at most J + 1; D:973-977, 3841)
    allocate(blk(jrot + 1))
    nfile = 0
    do while (nfile < jrot + 1)
      block
        double precision :: xtmp(npntt)
        integer :: rec4(4)
        read(iwave, iostat=ios) xtmp
        if (ios /= 0) exit
        nfile = nfile + 1
        associate (b => blk(nfile))
          read(iwave) rec4
          b%kz = rec4(1); b%maxleg = rec4(2); b%idvr = rec4(3); b%lincr = rec4(4)
          allocate(b%xalf(b%idvr), b%pleg(0:b%maxleg, b%idvr))
          b%xalf = xtmp(1:b%idvr)
          read(iwave) b%pleg
          read(iwave) b%meval
          allocate(b%eval(b%meval), b%vecs(b%idvr * npnt1 * npnt2, b%meval))
          read(iwave) b%eval
          do i = 1, b%meval
            read(iwave) b%vecs(:, i)
          end do
        end associate
      end block
    end do
    close(iwave)

    read(5, '(12i5)') nvib, neval, kmin, ibass, neval2
    mbass = npnt1 * npnt2 * npntt
    nvib = min(nvib, meval)
    nblk = jrot
    if (kmin > 0) nblk = nblk + 1
    nbass = nblk * nvib
    if (ibass > 0) nbass = min(nbass, ibass)
    if (neval <= 0) neval = 10
    neval = min(neval, nbass)
    if (neval2 <= 0) neval2 = neval
    write(6, 1001) meval, mbass, nvib, neval, nbass
    if (ibass > 0) write(6, 1005)
    read(5, '(9a8)') title
    write(6, 1010) title
    if (.not. zdiag .or. ires /= 0 .or. z1da) then
      write(6, '(/5x,a)') 'restart (ires /= 0), Hamiltonian-only (zdiag = F) and DVR1DA runs'
      write(6, '(5x,a)') 'are not supported by this build'
      stop 1
    end if
    if (zdcore) then
      write(6, 1013)
    else
      write(6, 1014)
      if (toler /= x0) write(6, 1060) toler
      if (toler == x0) write(6, 1070)
    end if
    if (zpham) write(6, 1020)
    if (.not. zpham) write(6, 1030)
    if (zpvec) write(6, 1040) thresh
    if (.not. zpvec) write(6, 1050)
    if (ztran) then
      zvec = .true.
      if (zptra) write(6, 1052)
      if (.not. zptra) write(6, 1053)
    end if
    write(6, 1051) iwave
    if (zpfun) write(6, 1055) ilev
    if (zvec) write(6, 1054) jvec
    if (zvec .and. kmin > 1) write(6, 1056) jvec2
    if (ztran) write(6, 1057) jscr
    if (ztran) write(6, 1058) kvec
    if (ztran .and. kmin > 1) write(6, 1059) kvec2
    if (toler /= x0) write(6, 1066) toler
    if (toler == x0) write(6, 1077)
    if (idia /= 2) then
      if (idia == 1) write(6, 1080)
      if (idia == 0) write(6, 1085)
    else
      if (ipar == 0) write(6, 1100)
      if (ipar == 1) write(6, 1110)
    end if
    write(6, 1120) jrot
    if (kmin == 0) write(6, 1130)
    if (kmin == 1) write(6, 1140)
    if (kmin > 1) write(6, 1150)
    if (kmin > 0 .and. kmin0 /= 1) then
      write(6, 970) kmin, kmin0
      stop
    end if
    if (zembed) then
      write(6, 1160) 2
    else
      write(6, 1160) 1
      zquad2 = .true.
    end if
    if (jrot == 1 .and. kmin == 0) then
      write(6, '(/5x,a)') 'J = 1 with KMIN = 0 is not supported by this build'
      stop 1
    end if
1001 format(/5x,'ROTATIONAL PART OF ROT-VIB CALCULATION  WITH:', &
            /i9,3x,'LOWEST VIBRATIONAL EIGENVECTORS SUPPLIED FROM', &
            /i9,3x,'DIMENSION VIBRATION SECULAR PROBLEM', &
            /i9,3x,'LOWEST VIBRATIONAL EIGENVECTORS ACTUALLY USED', &
            /i9,3x,'LOWEST ROTATIONAL EIGENVECTORS REQUIRED FOR', &
            /i9,3x,'DIMENSION ROTATION SECULAR PROBLEM')
1005 format(17x,'WITH BASIS SELECTED BY ENERGY ORDERING')
1010 format(/5x,'TITLE: ',9a8)
1013 format(/5x,'Diagonalisation performed in core using DSYEV')
1014 format(/5x,'Diagonalisation performed iteratively using F02FJF')
1060 format(5x,'EIGENVALUE CONVERGENCE TOLERANCE, TOLER =',d20.3)
1070 format(5x,'EIGENVALUES CONVERGED TO MACHINE ACCURACY')
1020 format(/5x,'PRINTING OF HAMILTONIAN MATRIX REQUESTED')
1030 format(/5x,'PRINTING OF HAMILTONIAN MATRIX NOT REQUESTED')
1040 format(5x,'PRINTING OF EIGENVECTOR COEFFICIENTS GREATER THAN',' THRESH =',f5.2,' REQUESTED')
1050 format(5x,'PRINTING OF EIGENVECTORS NOT REQUESTED')
1052 format(5x,'PRINTING OF TRANSFORMED VECTORS REQUESTED')
1053 format(5x,'PRINTING OF TRANSFORMED VECTORS NOT REQUESTED')
1051 format(/5x,'DVR3D   DATA   TO BE READ         FROM STREAM ','IWAVE =',i4)
1055 format(/5x,'EIGENVALUES    TO BE WRITTEN TO END OF STREAM ','ILEV  =',i4)
1054 format(/5x,'EIGENVALUES & VECTORS TO BE WRITTEN TO STREAM ','JVEC  =',i4)
1056 format( 5x,'SECOND SET            TO BE WRITTEN TO STREAM ','JVEC2 =',i4)
1057 format(/5x,'Transformed input vectors stored on    stream ','JSCR  =',i4)
1058 format(/5x,'TRANSFORMED VECTORS   TO BE WRITTEN TO STREAM ','KVEC  =',i4)
1059 format( 5x,'SECOND SET            TO BE WRITTEN TO STREAM ','KVEC2 =',i4)
1066 format(/5x,'EIGENVALUE CONVERGENCE TOLERANCE, TOLER =',d20.3)
1077 format(/5x,'EIGENVALUE CONVERGENCE TO MACHINE ACCURACY')
1080 format(/5x,'DIATOMIC ASSUMED HETRONUCLEAR')
1085 format(/5x,'RADAU COORDINATES USED')
1100 format(/5x,'DIATOMIC ASSUMED HOMONUCLEAR',/5x,'EVEN PARITY FUNCTIONS IN BASIS SET')
1110 format(/5x,'DIATOMIC ASSUMED HOMONUCLEAR',/5x,'ODD PARITY FUNCTIONS IN BASIS SET')
1120 format(/5x,'J =',i3,' ROTATIONAL STATE')
1130 format(12x,'WITH ANTI-SYMMETRIC |JK> - |J-K> FUNCTIONS IN BASIS')
1140 format(12x,'WITH SYMMETRIC |JK> + |J-K> FUNCTIONS IN BASIS')
1150 format(12x,'LOOP OVER SYMMETRIC AND ANTI-SYMMETRIC FUNCTIONS')
1160 format(/5x,'Z AXIS EMBEDED ALONG THE R',i1,' COORDINATE')
970 format(//6x,'**** KMIN =',i3,' BUT KMIN0 =',i3,' STOP ****',/)
  end subroutine insize

  ! SELECT (R:301-478): the blocks used (k = 0 skipped for an f-only run),
  ! MVIB per block, and the IBASS lowest functions over all blocks when
  ! IBASS asks for fewer than are there (on equal energies the lower k
  ! block wins)
  subroutine select_basis()
    integer, allocatable :: use(:), iv(:)
    integer :: nb, ipt, jpt, n, j, ivib, ipd, ipu, kz
    double precision :: evibr, emin
    allocate(use(nblk), mvib(nblk), iv(nblk), jst(nblk), nang(nblk))
    nb = 0
    do i = 1, nfile
      if (blk(i)%kz == 0 .and. kmin == 0) cycle
      nb = nb + 1
      if (nb > nblk) exit
      use(nb) = i
    end do
    if (nb < nblk) then
      write(6, '(/,a,i4)') '   END OF STREAM iwave = ', iwave
      stop
    end if
    blk = blk(use)
    kbass = 0
    ipt = 0
    do k = 1, nblk
      mvib(k) = min(nvib, blk(k)%meval)
      ipt = ipt + mvib(k)
      kbass = kbass + blk(k)%idvr * npnt1 * npnt2
    end do
    if (ibass <= 0 .or. ibass >= ipt) then
      ibass = ipt
      ivib = nvib
    else
      emin = blk(1)%eval(1)
      do k = 1, nblk
        emin = min(emin, blk(k)%eval(1))
        iv(k) = 1
      end do
      ipt = 1
      do n = 1, ibass
        do while (iv(ipt) > mvib(ipt))
          ipt = ipt + 1
        end do
        evibr = blk(ipt)%eval(iv(ipt))
        jpt = ipt
        do j = ipt + 1, nblk
          if (iv(j) > mvib(j)) cycle
          if (blk(j)%eval(iv(j)) >= evibr) cycle
          evibr = blk(j)%eval(iv(j))
          jpt = j
        end do
        iv(jpt) = iv(jpt) + 1
      end do
      mvib = iv - 1
      ivib = maxval(mvib)
      write(6, 1000) nbass, emin, evibr
    end if
    write(6, 1010)
    kz = -1
    if (kmin == 0) kz = 0
    ipu = 0
    do j = 1, nblk
      kz = kz + 1
      ipd = ipu
      ipu = ipd + mvib(j)
      if (mvib(j) == nvib) then
        write(6, 1021) kz, ipd + 1, ipu, 1.d2 * dble(mvib(j)) / dble(nvib)
      else
        write(6, 1020) kz, ipd + 1, ipu, 1.d2 * dble(mvib(j)) / dble(nvib)
      end if
    end do
    nbass = ibass
    nvib = ivib
1000 format(/,i9,'FUNCTIONS SELECTED FROM E =',d20.10,' TO',d20.10)
1010 format(//5x,' BASIS FUNCTIONS SELECTED')
1020 format(5x,'K =',i3,', I RUNS FROM',i5,' TO',i5,' ',f6.2,' %')
1021 format(5x,'K =',i3,', I RUNS FROM',i5,' TO',i5,' ',f6.2,' % *saturated*')
  end subroutine select_basis

  subroutine vrmain()
    type(dvr_rot_dims_t) :: d
    type(dvr_rot_block_t), allocatable :: bd(:)
    double precision, allocatable :: pl(:), xv(:), ev(:), eval(:), vec(:), dvec(:)
    integer(c_int), allocatable :: jn(:)
    integer(c_int) :: rc
    integer(c_int64_t) :: np, nx, nv, nd, p, q
    integer :: nrad, nsets, ib2, nev2, set, k1
    logical :: j1f
    nrad = npnt1 * npnt2
    ib2 = nbass - mvib(1)
    nev2 = min(neval2, ib2)
    j1f = kmin > 1 .and. jrot == 1
    nsets = 1
    if (kmin > 1 .and. .not. j1f) nsets = 2
    allocate(bd(nblk), jn(2 * nblk))
    np = 0; nx = 0
    do k = 1, nblk
      bd(k)%maxleg = blk(k)%maxleg; bd(k)%idvr = blk(k)%idvr
      bd(k)%lincr = blk(k)%lincr; bd(k)%mvib = mvib(k)
      np = np + int(blk(k)%maxleg + 1, 8) * blk(k)%idvr
      nx = nx + int(mvib(k), 8) * blk(k)%idvr * nrad
      call angular(k, jst(k), nang(k))
    end do
    allocate(pl(np), xv(max(1_8, nx)), ev(nbass))
    p = 0; q = 0
    do k = 1, nblk
      pl(p + 1:p + (blk(k)%maxleg + 1) * blk(k)%idvr) = reshape(blk(k)%pleg, [size(blk(k)%pleg)])
      p = p + size(blk(k)%pleg)
      do i = 1, mvib(k)
        xv(q + 1:q + size(blk(k)%vecs, 1)) = blk(k)%vecs(:, i)
        q = q + size(blk(k)%vecs, 1)
      end do
    end do
    p = 0
    do k = 1, nblk
      ev(p + 1:p + mvib(k)) = blk(k)%eval(1:mvib(k))
      p = p + mvib(k)
    end do
    d%nblk = nblk; d%kfirst = 0
    if (kmin == 0) d%kfirst = 1
    d%jrot = jrot; d%idia = idia; d%ipar = ipar; d%npnt1 = npnt1; d%npnt2 = npnt2
    d%zembed = 0; if (zembed) d%zembed = 1
    d%zquad2 = 0; if (zquad2) d%zquad2 = 1
    d%nlim = nlim; d%nsets = nsets
    d%neval(1) = neval; d%neval(2) = 0
    if (nsets == 2) d%neval(2) = nev2
    d%nvec = 0
    if (zvec .or. ztran .or. zpvec) d%nvec = 1
    nv = int(nbass, 8) * neval
    nd = 0
    do k = 1, nblk
      nd = nd + int(nang(k), 8) * nrad * neval
    end do
    if (nsets == 2) then
      nv = nv + int(ib2, 8) * nev2
      do k = 2, nblk
        nd = nd + int(nang(k), 8) * nrad * nev2
      end do
    end if
    allocate(eval(neval + nev2), vec(max(1_8, nv)), dvec(max(1_8, nd)))
    rc = dvr_rotlev3(d, bd, pl, xv, ev, rm2, jn, eval, vec, dvec)
    if (rc /= DVR_OK) then
      write(6, '(5x,a,i4)') 'dvr_rotlev3 failed, code', rc
      stop 1
    end if
    write(6, '(/5x,a)') 'HAMILTONIAN CONSTRUCTION COMPLETE'
    ezero = x0
    read(5, '(f20.0)', iostat=ios) ezero
    if (ios /= 0) ezero = x0
    ! first set: all blocks (e parity for KMIN >= 1, f for KMIN = 0)
    call set_heading(kmin >= 1, nbass)
    call dgrot(1, nbass, neval, eval(1:neval), vec, jvec)
    write(6, '(/5x,a)') 'DIAGONALISATION COMPLETE'
    if (ztran) call dstore(1, neval, eval(1:neval), dvec, kvec)
    if (kmin <= 1) return
    ! second set (KMIN > 1): blocks k >= 1, f parity
    if (j1f) then
      call dstore1(nev2, ib2)
      return
    end if
    k1 = 2
    call set_heading(.false., ib2)
    call dgrot(k1, ib2, nev2, eval(neval + 1:neval + nev2), vec(nbass * neval + 1:), jvec2)
    write(6, '(/5x,a)') 'DIAGONALISATION COMPLETE'
    if (ztran) call dstore(2, nev2, eval(neval + 1:neval + nev2), &
                           dvec(nd_of_set1() + 1:), kvec2)
  end subroutine vrmain

  integer(c_int64_t) function nd_of_set1()
    nd_of_set1 = 0
    do k = 1, nblk
      nd_of_set1 = nd_of_set1 + int(nang(k), 8) * npnt1 * npnt2 * neval
    end do
  end function

  ! first l and count of the associated Legendre functions of block k
  ! (JTRAN R:879-885; SOLRT's K is the block's k)
  subroutine angular(k, jstart, na)
    integer, intent(in) :: k
    integer, intent(out) :: jstart, na
    integer :: kk, jdia
    jdia = max(idia, 1)
    kk = k - 1
    if (kmin == 0) kk = k
    jstart = kk
    if (mod(jstart, jdia) /= ipar) jstart = jstart + 1
    na = (blk(k)%maxleg + blk(k)%lincr - jstart) / jdia + 1
  end subroutine

  subroutine set_heading(eparity, nb)
    logical, intent(in) :: eparity
    integer, intent(in) :: nb
    if (.not. eparity) write(6, 1000) jrot, nb
    if (eparity) write(6, 1010) jrot, nb
    if (idia == 2) then
      if (ipar == 0) write(6, 1020)
      if (ipar == 1) write(6, 1030)
    end if
1000 format(5x,'J =',i3,' ROTATIONAL STATE,',i7,' BASIS FUNCTIONS', &
            /12x,'F PARITY, ANTI-SYMMETRIC |JK> - |J-K> FUNCTIONS IN BASIS')
1010 format(5x,'J =',i3,' ROTATIONAL STATE,',i7,' BASIS FUNCTIONS', &
            /12x,'E PARITY, SYMMETRIC |JK> + |J-K> FUNCTIONS IN BASIS')
1020 format(12x,'EVEN PARITY FUNCTIONS IN BASIS SET')
1030 format(12x,'ODD PARITY FUNCTIONS IN BASIS SET')
  end subroutine

  ! the outputs of DGROT (R:1046-1197) for set k1 (1 or 2): stdout, ILEV,
  ! JVEC, fort.62 / fort.63
  subroutine dgrot(k1, nb, nev, eval, vec, junit)
    integer, intent(in) :: k1, nb, nev, junit
    double precision, intent(in) :: eval(nev), vec(nb, nev)
    double precision :: evalcm(nev), vv, vbig
    integer :: ip, kz, mbeg, mend, kk, j, ibig(nb), ipt
    if (zdcore) write(6, *) ' Found ', nev, ' evalues (out of ', nev, ')'
    write(6, 1000) nev
    write(6, 1020) eval
    ip = 1 - kmin
    if (kmin > 1) ip = k1 - 1
    if (zpfun) then
      open(unit=ilev, form='formatted', position='append')
      write(ilev, 1025) jrot, ip, idia, ipar, 0, nev, nb
      write(ilev, 1026) eval
      close(ilev)
    end if
    if (zvec) then
      kz = kmin
      if (kmin > 1) kz = 2 - k1
      open(unit=junit, form='unformatted')
      rewind junit
      write(junit) jrot, kz, ipar, nev, nb
      write(junit) (mvib(kk), kk = k1, nblk)
      write(junit) eval
      mend = 0
      do kk = k1, nblk
        mbeg = mend + 1
        mend = mend + mvib(kk)
        if (mvib(kk) > 0) write(junit) ((vec(j, i), j = mbeg, mend), i = 1, nev)
      end do
      close(junit)
    end if
    do i = 1, nev
      evalcm(i) = eval(i) * autocm - ezero
    end do
    write(6, 1010) nev, ezero
    write(6, 1022) evalcm
    do i = 1, nev
      if (ip == 0) then
        write(62, 1034) i, eval(i), evalcm(i)
      else
        write(63, 1034) i, eval(i), evalcm(i)
      end if
    end do
    if (zpvec) then
      if (thresh <= x0) then
        write(6, '(//,a,/)') '    EIGENVECTORS'
        do i = 1, nev
          write(6, '(/(1x,10f13.7))') (vec(j, i), j = 1, nb)
        end do
      else
        write(6, 1050) thresh
        do i = 1, nev
          vbig = x0
          ipt = 0
          do j = 1, nb
            vv = abs(vec(j, i))
            if (vv > thresh) then
              ipt = ipt + 1
              ibig(ipt) = j
            end if
            if (ipt <= 0 .and. vv > vbig) then
              vbig = vv
              ibig(1) = j
            end if
          end do
          write(6, 1060) i, (ibig(j), vec(ibig(j), i), j = 1, max(1, ipt))
        end do
      end if
    end if
1000 format(//5x,'LOWEST',i4,' EIGENVALUES IN HARTREES',/)
1020 format(5d24.12)
1025 format(7i4)
1026 format(4d20.12)
1010 format(//5x,'Lowest',i4,' eigenvalues in wavenumbers relative to',' EZERO =',d24.12/)
1022 format(1x,10f13.5/)
1034 format(1i4,2f30.20)
1050 format(//,'EIGENVECTOR COMPONANTS GREATER THAN THRESH =',f5.2)
1060 format('  VECTOR',i3,5(i7,f14.10)/(11x,5(i7,f14.10)))
  end subroutine dgrot

  ! DSTORE (R:1201-1341): the vectors of set itra in the FBR basis for
  ! DIPOLE3, one record per k block of the set.  Each record is the first
  ! nkbas*nval values of the (mbass, nval) array the reference fills with
  ! dgemm (ldd = mbass): packed when nkbas = mbass.
  subroutine dstore(itra, nval, energy, dv, kunit)
    integer, intent(in) :: itra, nval, kunit
    double precision, intent(in) :: energy(nval), dv(*)
    integer :: lblk, kz, nend, kk, nk, j
    integer :: lmin(nblk), lbasis(nblk), nkbas(nblk)
    integer(c_int64_t) :: p
    double precision, allocatable :: rec(:)
    write(6, 1000) iwave, jscr, merge(jvec, jvec2, itra == 1), kunit
    lblk = nblk - itra + 1
    kz = kmin
    if (kmin > 1) kz = 2 - itra
    nend = 0
    do kk = 1, lblk
      lmin(kk) = jst(kk + itra - 1)
      lbasis(kk) = nang(kk + itra - 1)
      nkbas(kk) = nang(kk + itra - 1) * npnt1 * npnt2
      nend = nend + nkbas(kk)
    end do
    if (nend > kbass) then
      write(6, 930) nend, kbass
      stop
    end if
    open(unit=kunit, form='unformatted')
    write(kunit) idia, ipar, npntt, npnt1, npnt2, jrot, kz, nval
    write(kunit) zembed, zmors1, zmors2, xmass, g1, g2, .false.
    write(kunit) morse
    write(kunit) nend, lmin(1:lblk), lbasis(1:lblk), nkbas(1:lblk)
    write(kunit) r1
    write(kunit) r2
    write(kunit) nval
    write(kunit) energy
    p = 0
    do kk = 1, lblk
      nk = nkbas(kk)
      allocate(rec(int(nk, 8) * nval))
      rec = x0
      do i = 1, nval
        do j = 1, nk
          if (int(i - 1, 8) * mbass + j <= int(nk, 8) * nval) &
              rec(int(i - 1, 8) * mbass + j) = dv(p + int(i - 1, 8) * nk + j)
        end do
      end do
      write(kunit) rec
      if (zptra) then
        write(6, '(/a,i3/)') 'Vectors for K-block', kk
        do i = 1, nval
          write(6, '(1x,10f13.7)') (dv(p + int(i - 1, 8) * nk + j), j = 1, nk)
        end do
      end if
      deallocate(rec)
      p = p + int(nk, 8) * nval
    end do
    close(kunit)
    write(6, 1050) jrot, kz, ipar, nval, nend
1000 format(5x,'EIGENVECTOR TRANSFORMATION:', &
            /5x,'INPUT:  DVR3D data,            IWAVE =',i3, &
            /5x,'        ROTLEV3 data (scratch) JSCR  =',i3, &
            /5x,'        ROTLEV3 data (final)   JVEC  =',i3, &
            /5x,'OUTPUT: transformed vectors,   KVEC  =',i3/)
930 format(/5x,'ERROR IN SIZE OF TOTAL BASIS',/5x,'NEND =',i7,'  KBASS =',i7/)
1050 format(//5x,'TRANSFORMATION COMPLETED SUCCESSFULLY ', &
            'FOR STATE: JROT =',i3,', KMIN =',i3,', IPAR =',i3, &
            /5x,'STORED',i4,' VIB-ROT LEVELS',5x,'NEND =',i7,/)
  end subroutine dstore

  ! DSTORE1 (R:1344-1475): the f set of J = 1 is DVR3DRJZ's own k = 1 block
  subroutine dstore1(nev, ib)
    integer, intent(in) :: nev, ib
    double precision :: ev(nev)
    integer :: nk
    ev = blk(2)%eval(1:nev)
    nk = nang(2) * npnt1 * npnt2
    if (ztran) then
      open(unit=kvec2, form='unformatted')
      write(kvec2) idia, ipar, npntt, npnt1, npnt2, jrot, 0, nev
      write(kvec2) zembed, zmors1, zmors2, xmass, g1, g2, .false.
      write(kvec2) morse
      write(kvec2) nk, jst(2), nang(2), nk
      write(kvec2) r1
      write(kvec2) r2
      write(kvec2) blk(2)%xalf
      write(kvec2) blk(2)%kz, blk(2)%maxleg, blk(2)%idvr, blk(2)%lincr
      write(kvec2) blk(2)%pleg
      write(kvec2) nev
      write(kvec2) ev
    end if
    write(6, 1000) nev
    write(6, 1020) ev
    if (zpfun) then
      open(unit=ilev, form='formatted', position='append')
      write(ilev, '(7i4)') jrot, 1, idia, ipar, 0, nev, ib
      write(ilev, '(4d20.12)') ev
      close(ilev)
    end if
    write(6, 1010) nev, ezero
    write(6, 1022) (ev(i) * autocm - ezero, i = 1, nev)
    if (.not. ztran) return
    write(6, 2000) iwave, jscr, kvec2
    if (zptra) write(6, 2010) jrot, 0, ipar, nev, 1
    do i = 1, nev
      write(kvec2) blk(2)%vecs(1:nk, i)
      if (zptra) write(6, '(1x,10f13.7)') blk(2)%vecs(1:nk, i)
    end do
    write(6, 2050) jrot, 0, ipar, nev, nk
    close(kvec2)
1000 format(//5x,'LOWEST',i4,' EIGENVALUES IN HARTREES',/)
1020 format(5d24.12)
1010 format(//5x,'Lowest',i4,' eigenvalues in wavenumbers relative to',' EZERO =',d24.12/)
1022 format(1x,10f13.5/)
2000 format(5x,'EIGENVECTOR DUMMY TRANSFORMATION J=1f case:', &
            /5x,'INPUT:  DVR3D data,            IWAVE =',i3, &
            /5x,'        ROTLEV3 data (scratch) JSCR  =',i3, &
            /5x,'OUTPUT: transformed vectors,   KVEC  =',i3/)
2010 format(/5x,'J =',i3,' KMIN =',i2,'  PARITY =',i2,' NVAL =',i3/ &
            /'Vectors for K-block',i3/)
2050 format(//5x,'Transfer of vectors completed successfully', &
            //5x,'for rotational state: jrot =',i3,', kmin =',i3, &
            ', ipar =',i3,'.  stored',i6,' vib-rot levels', &
            /5x,'nend =',i7)
  end subroutine dstore1

end program rotlev3_amd
