! dvr3drjz_amd -- drop-in replacement for the DVR3DRJZ executable on the
! theta-first jhmain path (J = 0; J > 0 at a fixed k with zrot = F; J > 0 with
! zrot = T over the k blocks of ccmain), running the Hamiltonian stages on one
! MI355X through libdvr3d_amd.
!
! Same contract as the reference program (source/dvr3drjz_segmented.f90):
!   stdin   the .job deck (namelist &PRT + cards, docum.tex:72-235)
!   link    the user's SUBROUTINE POTV(V,R1,R2,XCOS) with COMMON /MASS/
!   stdout  the eigenvalue report of diag3d per k block (D:2703-2761 formats)
!   ILEV    formatted levels when zpfun (D:469-471, D:2708-2719)
!   IWAVE   unformatted wavefunction file when ztran (records of setcon
!           D:1162-1180, ccmain D:820-845 and D:973-977, trans D:3733-3841),
!           in the compiler's default sequential record format.
!
! The k blocks of a zrot run (ccmain loop, D:857-1007) are independent: every
! block gets its own device context, all of them run concurrently
! (dvr_run_blocks), and the reports and IWAVE records are written afterwards
! in k order, as the reference writes them.
program dvr3drjz_amd
  use iso_c_binding
  use dvrh_setup
  use dvr3d_amd_api
  implicit none
  type(dvrh_job_t) :: job
  type(dvrh_sizes_t) :: sz
  type(dvrh_consts_t) :: cst
  type(dvr_dims_t) :: dims
  type(c_ptr), allocatable :: ctx(:)
  real(c_double), allocatable :: r1(:), r2(:), xk1(:, :), xk2(:, :), xalf(:, :), walf(:, :)
  real(c_double), allocatable :: pleg(:, :, :), xl(:, :), wsum(:), wterm(:), vgrid(:, :, :)
  real(c_double), allocatable :: xla(:, :), eval(:), evalcm(:), psi(:), xkb(:, :), xkc(:, :)
  real(c_double) :: xmass(3), g1, g2, xmassr(3), v, ezero, ms(7), mst(7)
  common /mass/ xmass, g1, g2, xmassr
  character(len=72) :: title
  integer :: ierr, i, a, b, c, kz, ik, nk, meval, ii, jdia, jpar, isym, ip, jstart, nang
  integer :: mbass, nlim, lincr
  integer(c_int) :: rc
  integer(8) :: t0, t1, trate
  logical :: zpfun, ztran, zembed, zmors1, zmors2, zncor, zquad2

  call system_clock(t0, trate)
  write(6, '(5x,a)') 'Program DVR3DRJZ (MI355X / gfx950 build of the theta-first path)'
  call dvrh_read_job_unit(5, job, ierr)
  if (ierr /= 0) then
    write(6, '(5x,a,i4)') 'error reading the .job input, code', ierr
    stop 1
  end if
  if (job%idia == -2) then
    ! symmetrised Radau coordinates: the nfmain driver (D:988-1003)
    call run_nf()
    call system_clock(t1)
    write(6, '(/i10,a/)') int((t1 - t0) / trate), ' secs CPU time used'
    stop
  end if
  ierr = dvrh_sizes(job, sz)
  if (ierr /= 0) then
    write(6, '(5x,a,i4)') 'option outside the theta-first jhmain path, code', ierr
    stop 1
  end if
  do i = 1, 72
    title(i:i) = job%title(i)
  end do
  write(6, 1030) job%npnt1, job%npnt2, job%nalf, sz%neval, sz%max3d
1030 format(/5x,i5,3x,'radial r1 dvr points used,'/5x,i5,3x,'radial r2 dvr points used,', &
            /5x,i5,3x,'angular dvr points used, with',/5x,i5,3x,'lowest eigenvectors chosen from', &
            /5x,i5,3x,'maximum dimension secular problem'/)
  write(6, '(5x,a,a)') 'Title: ', title
  nk = sz%ku - sz%kd + 1
  ! with zladd = F the angular basis shrinks with k (D:930-950): not built
  if (nk > 1 .and. job%zladd == 0) then
    write(6, '(5x,a)') 'zrot k-loop with zladd = F is not available in this build'
    stop 1
  end if
  if (sz%jrot > 0 .and. sz%zrot /= 0) write(6, '(/5x,a/5x,a,i4,a,i3,a,i3)') &
      '***  vibrational part of rot-vib calculation  ***', 'J =', sz%jrot, ',  k =', sz%kd, &
      ' ..', sz%ku

  xmass = job%xmass; xmassr = job%xmassr
  ztran = job%ztran /= 0
  zpfun = job%zpfun /= 0
  if (abs(job%jrot) > 1) zpfun = .false.        ! D:466
  allocate(r1(job%npnt1), r2(job%npnt2), xk1(job%npnt1, job%npnt1), xk2(job%npnt2, job%npnt2))
  allocate(xalf(sz%idvr, nk), walf(sz%idvr, nk), pleg(0:sz%maxleg, sz%idvr, nk))
  allocate(xl(sz%idvr, sz%idvr), wsum(sz%npntb * sz%npntc), wterm(sz%npntb * sz%npntc))
  allocate(vgrid(sz%npnta, sz%npntb, sz%npntc))
  allocate(xla(sz%npnta, sz%npnta), xkb(sz%npntb, sz%npntb), xkc(sz%npntc, sz%npntc))
  allocate(ctx(nk))
  ctx = c_null_ptr

  dims%npnta = sz%npnta; dims%npntb = sz%npntb; dims%npntc = sz%npntc
  dims%max2d = sz%max2d; dims%max3d = sz%max3d; dims%neval = sz%neval
  dims%zcut = job%zcut; dims%zall = job%zall; dims%zr2r1 = job%zr2r1
  dims%idvr = sz%idvr; dims%npnt1 = job%npnt1; dims%npnt2 = job%npnt2
  dims%nvec = 0
  if (ztran) dims%nvec = sz%neval
  dims%solver = DVR_SOLVER_DENSE

  ! ---- one context per k block: basis, rotational term, potential
  do ik = 1, nk
    kz = sz%kd + ik - 1
    ierr = dvrh_setup_basis(job, sz, kz, r1, r2, xk1, xk2, xalf(:, ik), walf(:, ik), &
                            pleg(:, :, ik), xl, cst)
    if (ierr /= 0) then
      write(6, '(//5x,a,i4//)') 'points & weights in error, adjust algorithm; code', ierr
      stop 1
    end if
    g1 = cst%g1; g2 = cst%g2
    ierr = dvrh_wsum(job, sz, cst, kz, r1, r2, wsum, wterm)
    ! potential on the grid through the user's POTV (mkham1, D:2178)
    do c = 1, sz%npntc
      do b = 1, sz%npntb
        do a = 1, sz%npnta
          if (job%zr2r1 /= 0) then
            call potv(v, r1(c), r2(b), xalf(a, ik))
          else
            call potv(v, r1(b), r2(c), xalf(a, ik))
          end if
          vgrid(a, b, c) = v
        end do
      end do
    end do
    rc = dvr_ctx_create(dims, 0_c_int, ctx(ik))
    if (rc /= DVR_OK) then
      write(6, '(5x,a,i4)') 'dvr_ctx_create failed, code', rc
      stop 1
    end if
    xla = xl(1:sz%npnta, 1:sz%npnta)
    if (job%zr2r1 /= 0) then
      xkb = xk2; xkc = xk1
    else
      xkb = xk1; xkc = xk2
    end if
    rc = dvr_set_basis(ctx(ik), xla, xkb, xkc, wsum, wterm, job%emax1 / autocm, &
                       job%emax2 / autocm)
    if (rc == DVR_OK) rc = dvr_set_potential(ctx(ik), vgrid)
    if (rc /= DVR_OK) then
      write(6, '(5x,a,i4,2x,a)') 'device set-up failed, code', rc, trim(dvr_error_text(ctx(ik)))
      stop 1
    end if
  end do

  rc = dvr_run_blocks(ctx, int(nk, c_int))
  if (rc /= DVR_OK) then
    do ik = 1, nk
      if (len_trim(dvr_error_text(ctx(ik))) == 0) cycle
      if (rc == DVR_ERR_CORE_2D .or. rc == DVR_ERR_CORE_3D) then
        write(6, '(//6x,a)') trim(dvr_error_text(ctx(ik)))
        stop
      end if
      write(6, '(5x,a,i3,a,i4,2x,a)') 'device pipeline failed at k =', sz%kd + ik - 1, &
          ', code', rc, trim(dvr_error_text(ctx(ik)))
      stop 1
    end do
    write(6, '(5x,a,i4)') 'device pipeline failed, code', rc
    stop 1
  end if

  ! ---- ILEV (opened once, D:467-478), IWAVE set-up records (D:1162-1180, D:820-845)
  if (zpfun) then
    open(unit=job%ilev, form='formatted')
    if (sz%jrot == 0 .and. mod(sz%ipar, 2) == 0) then
      write(job%ilev, '(9a8)') title
    else
      do
        read(job%ilev, *, iostat=ierr)
        if (ierr /= 0) exit
      end do
    end if
  end if
  if (ztran) then
    zembed = sz%zembed /= 0; zmors1 = job%zmors1 /= 0; zmors2 = job%zmors2 /= 0
    zncor = sz%zrot == 0; zquad2 = job%zquad2 /= 0
    nlim = 0
    if (sz%zrot /= 0) then
      if (zembed) then
        nlim = job%npnt2
      else
        nlim = job%npnt1
      end if
    end if
    open(unit=job%iwave, form='unformatted')
    write(job%iwave) sz%idia, sz%ipar, sz%idvr, job%npnt1, job%npnt2, sz%jrot, sz%kmin, &
                     sz%neval, nlim
    write(job%iwave) zembed, zmors1, zmors2, xmass, g1, g2, zncor, zquad2
    write(job%iwave) job%re1, job%diss1, job%we1, job%re2, job%diss2, job%we2
    if (nk > 1) then
      ! r**(-2) term for the second variational step (D:823-833)
      if (zembed) then
        write(job%iwave) (0.5d0 / (r2(i) * r2(i) * cst%urr2), i = 1, job%npnt2)
      else
        write(job%iwave) (0.5d0 / (r1(i) * r1(i) * cst%urr1), i = 1, job%npnt1)
      end if
    else
      jdia = max(1, sz%idia)
      jstart = sz%kmin
      if (mod(jstart, jdia) /= sz%ipar) jstart = jstart + 1
      nang = (sz%maxleg - jstart) / jdia + 1
      mbass = sz%idvr * job%npnt1 * job%npnt2
      write(job%iwave) mbass, jstart, nang, mbass
    end if
    write(job%iwave) r1
    write(job%iwave) r2
  end if

  ! ---- per k block: report (diag3d, D:2703-2761) and IWAVE records
  allocate(eval(sz%neval), evalcm(sz%neval))
  mst = 0.0d0
  do ik = 1, nk
    kz = sz%kd + ik - 1
    if (sz%zrot /= 0) write(6, '(/5x,a,i3,a,i3)') 'Solutions with J =', sz%jrot, ' k =', kz
    rc = dvr_get_eigenvalues(ctx(ik), eval, meval)
    if (job%zpmin == 0) then
      write(6, 1010) meval
      write(6, 1020) (eval(i), i = 1, meval)
    end if
    if (zpfun) then
      ip = sz%jrot - sz%kmin
      if (sz%jrot == 0 .or. ip == 1) then
        jdia = max(0, sz%idia)
        jpar = min(jdia, sz%ipar)
        isym = abs(min(0, sz%idia))
        if (sz%ipar == 1) isym = -isym
        write(job%ilev, '(6i4)') sz%jrot, ip, jdia, jpar, isym, meval
        write(job%ilev, '(4d20.12)') (eval(i), i = 1, meval)
      end if
    end if
    do i = 1, meval
      evalcm(i) = eval(i) * autocm
    end do
    write(6, 1031) meval
    write(6, 1020) (evalcm(i), i = 1, meval)
    if (sz%jrot == 0) then
      if (abs(sz%idia) == 2 .and. sz%ipar == 1) then
        ii = 1
        ezero = job%ezero
      else
        ezero = evalcm(1)
        ii = 2
      end if
      write(6, 1070)
      write(6, 1021) (evalcm(i) - ezero, i = ii, meval)
    end if
    if (ztran) then
      ! angular records of ccmain (D:973-977) then trans (D:3733-3841)
      lincr = kz
      write(job%iwave) xalf(:, ik)
      write(job%iwave) kz, sz%maxleg, sz%idvr, lincr
      write(job%iwave) ((pleg(i, a, ik) * walf(a, ik), i = 0, sz%maxleg), a = 1, sz%idvr)
      write(job%iwave) meval
      write(job%iwave) (eval(i), i = 1, meval)
      mbass = sz%idvr * job%npnt1 * job%npnt2
      if (.not. allocated(psi)) allocate(psi(int(mbass, 8) * sz%neval))
      rc = dvr_get_wavefunctions(ctx(ik), psi)
      do i = 1, meval
        write(job%iwave) psi(int(i - 1, 8) * mbass + 1:int(i, 8) * mbass)
      end do
    end if
    rc = dvr_get_stage_times(ctx(ik), ms)
    mst = mst + ms
  end do
1010 format(/5x,'lowest',i4,' eigenvalues in hartrees:',/)
1031 format(//5x,'Lowest',i4,' eigenvalues in wavenumbers:'/)
1020 format(5d24.12/)
1021 format(5f15.6/)
1070 format(//'  Band origins in wavenumbers:'/)
  if (zpfun) close(job%ilev)
  if (ztran) close(job%iwave)
  write(6, '(/5x,a,7f10.3)') 'device stage times (ms):', mst
  do ik = 1, nk
    call dvr_ctx_destroy(ctx(ik))
  end do
  call system_clock(t1)
  write(6, '(/i10,a/)') int((t1 - t0) / trate), ' secs CPU time used'

contains

  ! IDIA = -2, J = 0: set-up on the host, the potential through POTV on the
  ! (r_a, r_b, x_g) grid, everything else in one dvr_nfmain call; stdout,
  ! ILEV and IWAVE as nfmain, diag3d, testiv and transr write them
  ! (D:3066-3122, 2703-2761, 3447-3448, 3623-3641; set-up records D:820-845).
  subroutine run_nf()
    type(dvr_nf_dims_t) :: nd
    real(c_double), allocatable :: rr(:), hr(:, :), xa(:), wa(:), pl(:, :), ht(:, :), vg(:, :, :)
    real(c_double), allocatable :: wv(:)
    integer(c_int), allocatable :: iv2(:)
    integer(c_int) :: n3, iang
    integer :: npnt, nalf, npair, g, nbass
    npnt = job%npnt2; nalf = job%nalf
    allocate(rr(npnt), hr(npnt, npnt), xa(nalf), wa(nalf), pl(0:nalf - 1, nalf), ht(nalf, nalf))
    ierr = dvrh_setup_nf(job, sz, rr, hr, xa, wa, pl, ht, cst)
    if (ierr == 10) then
      write(6, '(5x,a)') 'IDIA = -2 is available for J = 0 with Morse radial bases in this build'
      stop 1
    else if (ierr /= 0) then
      write(6, '(//5x,a,i4//)') 'points & weights in error, adjust algorithm; code', ierr
      stop 1
    end if
    do i = 1, 72
      title(i:i) = job%title(i)
    end do
    write(6, 1030) npnt, npnt, nalf, sz%neval, sz%max3d
    write(6, '(5x,a,a)') 'Title: ', title
    xmass = job%xmass; xmassr = job%xmassr
    g1 = cst%g1; g2 = cst%g2
    ztran = job%ztran /= 0
    zpfun = job%zpfun /= 0
    npair = sz%max2d
    allocate(vg(npnt, npnt, nalf), iv2(nalf))
    do g = 1, nalf
      do b = 1, npnt
        do a = 1, npnt
          call potv(v, rr(a), rr(b), xa(g))
          vg(a, b, g) = v
        end do
      end do
    end do
    nd%npnt = npnt; nd%nalf = nalf; nd%ipar = sz%ipar
    nd%max3d = sz%max3d; nd%neval = sz%neval; nd%zcut = job%zcut
    nd%nvec = 0
    if (ztran) nd%nvec = sz%neval
    nd%jrot = 0; nd%kz = 0; nd%kmin = 0; nd%zrot = 0; nd%zlin = job%zlin
    allocate(eval(sz%neval), evalcm(sz%neval))
    allocate(wv(max(1, int(nd%nvec, 8) * npair * nalf)))
    write(6, '(5x,a)') 'starting the 2d calculations.'
    rc = dvr_nfmain(nd, hr, ht, rr, xa, cst%ur1, cst%urr1, vg, job%emax2 / autocm, iv2, n3, &
                    iang, eval, wv)
    if (rc == DVR_ERR_CORE_3D) then
      write(6, '(a,i4,a,e12.5,a)') 'number of 2d eigenvalues greater than ', sz%max3d, &
          ' increase max3d or reduce emax2 which is ', job%emax2, ' cm-1.'
      stop
    else if (rc /= DVR_OK) then
      write(6, '(5x,a,i4)') 'dvr_nfmain failed, code', rc
      stop 1
    end if
    do g = 1, nalf
      write(6, '(5x,a,i2,a,i3,a)') 'for gamma = ', g, ' selected ', iv2(g), ' energies.'
    end do
    write(6, '(25x,a,i5)') 'total = ', n3
    write(6, '(5x,a)') 'building the 3d hamiltonian.'
    write(6, '(5x,a)') 'diagonalising the 3d hamiltonian.'
    meval = min(sz%neval, n3)
    if (job%zpmin == 0) then
      write(6, 1010) meval
      write(6, 1020) (eval(i), i = 1, meval)
    end if
    if (zpfun) then
      ! ILEV (D:467-471, 2708-2719): title for even parity, then
      ! J, p, jdia = 0, jpar = 0, isym = +-2 and the levels
      open(unit=job%ilev, form='formatted')
      if (mod(sz%ipar, 2) == 0) then
        write(job%ilev, '(9a8)') title
      else
        do
          read(job%ilev, *, iostat=ierr)
          if (ierr /= 0) exit
        end do
      end if
      isym = 2
      if (sz%ipar == 1) isym = -2
      write(job%ilev, '(6i4)') 0, 0, 0, 0, isym, meval
      write(job%ilev, '(4d20.12)') (eval(i), i = 1, meval)
      close(job%ilev)
    end if
    do i = 1, meval
      evalcm(i) = eval(i) * autocm
    end do
    write(6, 1031) meval
    write(6, 1020) (evalcm(i), i = 1, meval)
    if (sz%ipar == 1) then
      ii = 1
      ezero = job%ezero
    else
      ezero = evalcm(1)
      ii = 2
    end if
    write(6, 1070)
    write(6, 1021) (evalcm(i) - ezero, i = ii, meval)
    if (ztran) then
      zembed = .true.; zmors1 = job%zmors1 /= 0; zmors2 = job%zmors2 /= 0
      zncor = .true.; zquad2 = job%zquad2 /= 0
      open(unit=job%iwave, form='unformatted')
      write(job%iwave) -2, sz%ipar, nalf, npnt, npnt, 0, 0, sz%neval, 0
      ! one radial basis for both bonds: its flags and parameters twice
      ! (D:1181-1184)
      write(job%iwave) zembed, zmors1, zmors1, xmass, g1, g2, zncor, zquad2
      write(job%iwave) job%re1, job%diss1, job%we1, job%re1, job%diss1, job%we1
      mbass = nalf * npair
      write(job%iwave) mbass, 0, nalf, mbass
      write(job%iwave) rr
      write(job%iwave) xa
      write(job%iwave) 0, nalf - 1, nalf, 0
      write(job%iwave) ((pl(i, a) * wa(a), i = 0, nalf - 1), a = 1, nalf)
      nbass = iang * npair
      write(job%iwave) iang, nbass
      write(job%iwave) (iv2(g), g = 1, nalf)
      write(job%iwave) sz%neval
      write(job%iwave) (eval(i), i = 1, sz%neval)
      do i = 1, sz%neval
        write(job%iwave) wv(int(i - 1, 8) * nbass + 1:int(i, 8) * nbass)
      end do
      close(job%iwave)
    end if
1010 format(/5x,'lowest',i4,' eigenvalues in hartrees:',/)
1031 format(//5x,'Lowest',i4,' eigenvalues in wavenumbers:'/)
1020 format(5d24.12/)
1021 format(5f15.6/)
1030 format(/5x,i5,3x,'radial r1 dvr points used,'/5x,i5,3x,'radial r2 dvr points used,', &
            /5x,i5,3x,'angular dvr points used, with',/5x,i5,3x,'lowest eigenvectors chosen from', &
            /5x,i5,3x,'maximum dimension secular problem'/)
1070 format(//'  Band origins in wavenumbers:'/)
  end subroutine run_nf

end program dvr3drjz_amd
