! Fortran host of the MI355X DVR3DRJZ path: .job input, basis set-up and
! the quantities the device kernels need.  Everything here is O(npnt^3) or
! less and runs once per k block on the CPU.
!
! Reference (cited as D:<line> = source/dvr3drjz_segmented.f90:<line>):
!   input      namelist &PRT D:54-65 + cards D:232-239, D:372, D:1042-1095
!   sizes      insize D:171-638
!   constants  setcon D:1016-1188
!   radial     setfac/norms/lagpt/laguer D:1191-1446, keints D:1495, k1k2 D:1603
!   angular    jacobi D:1635, allpts D:1747, asleg D:1775, lmatrx D:1846
! The quadratures and recurrences are the published ones the reference uses
! (Stroud & Secrest Gauss-Laguerre search with Newton polishing; Gauss-Jacobi
! roots by Newton on the three-term recurrence; normalised associated
! Legendre functions by upward recurrence), so the grids agree with the
! reference to rounding.
!
! C entry points (bind(C)) let the Python mirror and tests drive the same
! code as the drop-in executable dvr3drjz_amd.
module dvrh_setup
  use iso_c_binding
  implicit none
  private
  public :: dvrh_job_t, dvrh_sizes_t, dvrh_consts_t
  public :: dvrh_read_job_unit, dvrh_read_job, dvrh_sizes, dvrh_setup_basis, dvrh_wsum, dvrh_setup_nf
  public :: amtoau, autocm

  real(c_double), parameter :: amtoau = 1.8228883d03     ! D:1037
  real(c_double), parameter :: autocm = 2.19474624d+05   ! D:2681

  ! Every input item of the DVR3DRJZ deck (docum.tex:72-235).  Logicals are
  ! carried as 0/1 integers so the type is interoperable with C.
  type, bind(C) :: dvrh_job_t
    integer(c_int) :: ncoord, npnt1, npnt2, jrot, neval, nalf, max2d, max3d
    integer(c_int) :: idia, kmin, ipar, max3d2
    integer(c_int) :: zpham, zprad, zpvec, zrot, zladd, zembed, zmors1, zmors2
    integer(c_int) :: zpmin, zvec, zquad2, zdiag, zlmat, zcut, zall, zlin
    integer(c_int) :: zp1d, zp2d, zr2r1, ztheta, ztran, ztwod, zperp, zx, zs0, zs1, zpfun
    integer(c_int) :: idiag1, idiag2, iout1, iout2, iwave, ilev
    integer(c_int) :: has_ezero
    real(c_double) :: fixcos, xmass(3), xmassr(3), emax1, emax2
    real(c_double) :: re1, diss1, we1, re2, diss2, we2, ezero
    character(kind=c_char) :: title(72)
  end type

  ! Derived sizes after the insize rules (D:242-355).
  type, bind(C) :: dvrh_sizes_t
    integer(c_int) :: npnta, npntb, npntc, idvr, nalf2, maxleg, max2d, max3d, neval
    integer(c_int) :: jrot, kmin, idia, ipar, zrot, zembed, zvec, kd, ku
  end type

  type, bind(C) :: dvrh_consts_t
    real(c_double) :: g1, g2, ur1, ur2, urr1, urr2, beta1, beta2, a1, a2
    real(c_double) :: csx1, csa1, tsx1, tsa1, csx2, csa2, tsx2, tsa2, cswalf, tswalf
    integer(c_int) :: iu1, iu2
  end type

contains

  ! ------------------------------------------------------------------ input
  subroutine dvrh_read_job_unit(u, job, ierr)
    integer, intent(in) :: u
    type(dvrh_job_t), intent(out) :: job
    integer, intent(out) :: ierr
    logical :: zpham, zprad, zpvec, zrot, zladd, zembed, zmors2, zs0, zx, zs1, &
               zpmin, zvec, zquad2, zdiag, zlmat, zcut, zall, zlin, zp1d, zp2d, &
               zr2r1, ztheta, ztran, zmors1, ztwod, zperp, zpfun
    integer :: idiag1, idiag2, iout1, iout2, iwave, ilev, ieigs1, ivecs1, &
               ieigs2, ivecs2, ivint, iband, intvec
    namelist /prt/ zpham, zprad, zpvec, zrot, zladd, zembed, zmors2, zs0, zx, zs1, &
                   zpmin, zvec, zquad2, zdiag, zlmat, zcut, zall, zlin, &
                   zp1d, zp2d, zr2r1, ztheta, ztran, zmors1, ztwod, zperp, &
                   idiag1, idiag2, iout1, iout2, iwave, zpfun, ilev, &
                   ieigs1, ivecs1, ieigs2, ivecs2, ivint, iband, intvec
    character(len=72) :: title
    integer :: i, ios
    ! defaults of the reference's block data (D:153-167; the code, not the
    ! comments, is authoritative: zrot defaults to .true., zpfun to .false.)
    zpham = .false.; zprad = .false.; zpvec = .false.; zrot = .true.
    zladd = .true.; zembed = .true.; zmors2 = .true.; zpmin = .false.
    zvec = .false.; zquad2 = .true.; zcut = .false.; zdiag = .true.
    zlmat = .false.; zall = .false.; zp1d = .false.; zp2d = .false.
    zr2r1 = .true.; ztheta = .true.; zmors1 = .true.; ztran = .false.
    ztwod = .false.; zperp = .false.; zx = .false.; zs0 = .false.; zs1 = .false.
    zlin = .false.; zpfun = .false.
    ieigs1 = 7; ivecs1 = 3; ieigs2 = 2; ivecs2 = 4; ivint = 17; iband = 15
    intvec = 16; idiag1 = 20; idiag2 = 21; iout1 = 24; iout2 = 25; iwave = 26
    ilev = 14
    ierr = 0
    read(u, nml=prt, iostat=ios)
    if (ios /= 0) then
      ierr = 1; return
    end if
    read(u, '(15i5)', iostat=ios) job%ncoord
    if (ios /= 0) then
      ierr = 2; return
    end if
    read(u, '(15i5)', iostat=ios) job%npnt2, job%jrot, job%neval, job%nalf, job%max2d, &
         job%max3d, job%idia, job%kmin, job%npnt1, job%ipar, job%max3d2
    if (ios /= 0) then
      ierr = 3; return
    end if
    read(u, '(a72)', iostat=ios) title
    if (ios /= 0) title = ' '
    do i = 1, 72
      job%title(i) = title(i:i)
    end do
    read(u, '(3f20.0)', iostat=ios) job%fixcos
    read(u, '(3f20.0)', iostat=ios) job%xmass
    if (ios /= 0) then
      ierr = 4; return
    end if
    read(u, '(3f20.0)', iostat=ios) job%xmassr
    if (ios /= 0) job%xmassr = 0.0d0
    if (job%xmassr(1) <= 0.0d0) job%xmassr = job%xmass      ! D:1049
    read(u, '(3f20.0)', iostat=ios) job%emax1, job%emax2
    read(u, '(3f20.0)', iostat=ios) job%re1, job%diss1, job%we1
    read(u, '(3f20.0)', iostat=ios) job%re2, job%diss2, job%we2
    if (ios /= 0) then
      ierr = 5; return
    end if
    job%ezero = 0.0d0
    job%has_ezero = 0
    read(u, '(f20.0)', iostat=ios) job%ezero
    if (ios == 0) job%has_ezero = 1
    job%zpham = l2i(zpham); job%zprad = l2i(zprad); job%zpvec = l2i(zpvec)
    job%zrot = l2i(zrot); job%zladd = l2i(zladd); job%zembed = l2i(zembed)
    job%zmors1 = l2i(zmors1); job%zmors2 = l2i(zmors2); job%zpmin = l2i(zpmin)
    job%zvec = l2i(zvec); job%zquad2 = l2i(zquad2); job%zdiag = l2i(zdiag)
    job%zlmat = l2i(zlmat); job%zcut = l2i(zcut); job%zall = l2i(zall)
    job%zlin = l2i(zlin); job%zp1d = l2i(zp1d); job%zp2d = l2i(zp2d)
    job%zr2r1 = l2i(zr2r1); job%ztheta = l2i(ztheta); job%ztran = l2i(ztran)
    job%ztwod = l2i(ztwod); job%zperp = l2i(zperp); job%zx = l2i(zx)
    job%zs0 = l2i(zs0); job%zs1 = l2i(zs1); job%zpfun = l2i(zpfun)
    job%idiag1 = idiag1; job%idiag2 = idiag2; job%iout1 = iout1
    job%iout2 = iout2; job%iwave = iwave; job%ilev = ilev
  end subroutine

  integer(c_int) function l2i(z)
    logical, intent(in) :: z
    l2i = 0
    if (z) l2i = 1
  end function

  integer(c_int) function dvrh_read_job(path, job) bind(C, name="dvrh_read_job")
    character(kind=c_char), intent(in) :: path(*)
    type(dvrh_job_t), intent(out) :: job
    character(len=1024) :: fname
    integer :: i, u, ios, ierr
    fname = ' '
    do i = 1, 1024
      if (path(i) == c_null_char) exit
      fname(i:i) = path(i)
    end do
    open(newunit=u, file=trim(fname), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      dvrh_read_job = 100
      return
    end if
    call dvrh_read_job_unit(u, job, ierr)
    close(u)
    dvrh_read_job = ierr
  end function

  ! ------------------------------------------------------------------ sizes
  ! Returns 0, or an error code for options outside the theta-first jhmain
  ! path (idia=-2 nfmain, zperp, ztwod, ztheta=F, zquad2=F, spherical
  ! oscillators) or inconsistent input (D:533-541, D:618-637).
  integer(c_int) function dvrh_sizes(job, sz) bind(C, name="dvrh_sizes")
    type(dvrh_job_t), intent(in) :: job
    type(dvrh_sizes_t), intent(out) :: sz
    integer :: zrot, zembed, kmin
    dvrh_sizes = 0
    zrot = job%zrot; zembed = job%zembed; kmin = job%kmin
    if (job%jrot == 0) then                     ! D:242-246
      zembed = 1; kmin = 0; zrot = 0
    end if
    if (job%idia == -2 .or. job%idia == 0 .or. job%zperp /= 0 .or. job%ztwod /= 0) then
      dvrh_sizes = 10; return
    end if
    if (job%ztheta == 0) then
      dvrh_sizes = 11; return
    end if
    if (job%zquad2 == 0 .or. job%zmors1 == 0 .or. job%zmors2 == 0) then
      dvrh_sizes = 12; return
    end if
    if (job%ncoord /= 3) then
      dvrh_sizes = 13; return
    end if
    sz%nalf2 = (job%nalf + 1) / 2
    sz%maxleg = job%nalf - 1
    sz%npnta = job%nalf / max(1, job%idia)
    if (job%zr2r1 /= 0) then
      sz%npntb = job%npnt2; sz%npntc = job%npnt1
    else
      sz%npntb = job%npnt1; sz%npntc = job%npnt2
    end if
    if (job%zall /= 0) then
      sz%max2d = sz%npnta * sz%npntb
      sz%max3d = sz%max2d * sz%npntc
    else
      sz%max2d = min(job%max2d, sz%npnta * sz%npntb)
      sz%max3d = min(job%max3d, sz%npnta * sz%npntb * sz%npntc)
    end if
    sz%neval = job%neval
    if (sz%neval <= 0) sz%neval = 10
    sz%neval = min(sz%max3d, sz%neval)
    sz%ipar = job%ipar
    if (abs(job%idia) == 2) then
      if (job%ipar /= 0 .and. job%ipar /= 1) then
        dvrh_sizes = 14; return
      end if
      sz%idvr = sz%nalf2
      if (2 * sz%idvr /= job%nalf) then
        dvrh_sizes = 15; return
      end if
    else
      sz%idvr = job%nalf
      sz%ipar = 0
    end if
    sz%jrot = abs(job%jrot); sz%idia = job%idia
    if (zrot /= 0 .and. sz%jrot /= 0 .and. kmin /= 0) kmin = 1
    sz%kmin = kmin; sz%zrot = zrot; sz%zembed = zembed
    sz%zvec = job%zvec
    if (job%ztran /= 0) sz%zvec = 1
    if (zrot /= 0) then                          ! k loop bounds, D:806-815
      sz%kd = 1 - min(kmin, 1); sz%ku = sz%jrot
    else
      sz%kd = kmin; sz%ku = kmin
    end if
  end function

  ! --------------------------------------------------------------- constants
  subroutine set_consts(job, c)
    type(dvrh_job_t), intent(in) :: job
    type(dvrh_consts_t), intent(out) :: c
    real(c_double) :: a, b, xm(3), xr(3)
    xm = job%xmass; xr = job%xmassr
    if (job%idia >= 1) then                     ! scattering (Jacobi) coordinates
      c%g1 = xm(2) / (xm(2) + xm(3))
      c%g2 = 0.0d0
    else                                        ! Radau coordinates
      a = sqrt(xm(3) / (xm(1) + xm(2) + xm(3)))
      b = xm(2) / (xm(1) + xm(2))
      c%g1 = 1.0d0 - a / (a + b - a * b)
      c%g2 = 1.0d0 - a / (1.0d0 - b + a * b)
    end if
    c%ur1 = amtoau / (c%g2 * c%g2 / xm(1) + 1.0d0 / xm(2) + (1.0d0 - c%g2)**2 / xm(3))
    c%ur2 = amtoau / (1.0d0 / xm(1) + c%g1 * c%g1 / xm(2) + (1.0d0 - c%g1)**2 / xm(3))
    c%urr1 = amtoau / (c%g2 * c%g2 / xr(1) + 1.0d0 / xr(2) + (1.0d0 - c%g2)**2 / xr(3))
    c%urr2 = amtoau / (1.0d0 / xr(1) + c%g1 * c%g1 / xr(2) + (1.0d0 - c%g1)**2 / xr(3))
    ! Morse-like oscillator constants: beta = we sqrt(mu/2De), A = 4De/we
    c%beta1 = job%we1 * sqrt(0.5d0 * c%ur1 / job%diss1)
    c%a1 = 4.0d0 * job%diss1 / job%we1
    c%iu1 = int(c%a1 + 0.5d0)
    if (job%diss2 > 0.0d0 .and. job%we2 > 0.0d0) then
      c%beta2 = job%we2 * sqrt(0.5d0 * c%ur2 / job%diss2)
      c%a2 = 4.0d0 * job%diss2 / job%we2
      c%iu2 = int(c%a2 + 0.5d0)
    else
      ! no r2 basis card (IDIA = -2 decks use the r1 basis for both bonds)
      c%beta2 = 0.0d0; c%a2 = 0.0d0; c%iu2 = 0
    end if
  end subroutine

  ! ------------------------------------------------------------------ radial
  ! Monic Laguerre recurrence p_j = (x - b_j) p_{j-1} - c_j p_{j-2} with
  ! b_j = alf + 2j - 1, c_j = (j-1)(alf+j-1); returns p_n, p_n', p_{n-1}.
  subroutine lag_eval(x, n, alf, pn, dpn, pn1)
    real(c_double), intent(in) :: x, alf
    integer, intent(in) :: n
    real(c_double), intent(out) :: pn, dpn, pn1
    real(c_double) :: p0, p1, d0, d1, q, dq, bj, cj
    integer :: j
    p0 = 1.0d0; p1 = x - alf - 1.0d0
    d0 = 0.0d0; d1 = 1.0d0
    do j = 2, n
      bj = alf + dble(2 * j - 1)
      cj = dble(j - 1) * (alf + dble(j - 1))
      q = (x - bj) * p1 - cj * p0
      dq = (x - bj) * d1 - cj * d0 + p1
      p0 = p1; p1 = q; d0 = d1; d1 = dq
    end do
    pn = p1; dpn = d1; pn1 = p0
  end subroutine

  ! n-point Gauss-Laguerre rule for weight y^alf e^-y: points y (ascending) and
  ! pseudo-weights w (the weights divided by Gamma(n+alf+1), scaled by cc),
  ! plus the reference's consistency sums.
  subroutine gauss_laguerre(n, alf, cc, y, w, csx, csa, tsx)
    integer, intent(in) :: n
    real(c_double), intent(in) :: alf, cc
    real(c_double), intent(out) :: y(n), w(n), csx, csa, tsx
    real(c_double) :: x, xs, xprev, xprev2, step, p, dp, p1, pt, d, fi, r1, r2, ratio, fn
    integer :: i, it
    fn = dble(n)
    tsx = fn * (alf + fn)                 ! sum of the roots = n (alf + n)
    csx = 0.0d0; csa = 0.0d0
    ! first two roots by a forward sign scan, the rest from the Stroud &
    ! Secrest extrapolation; each polished by Newton to 1e-12 relative
    x = (1.0d0 + alf) * (2.0d0 + alf) / (1.0d0 + 3.0d0 * fn + 2.0d0 * alf)
    step = 3.0d0 * (1.0d0 + alf) / (1.0d0 + 3.0d0 * fn + alf)
    call lag_eval(x, n, alf, pt, dp, p1)
    xprev = 0.0d0; xprev2 = 0.0d0
    do i = 1, n
      if (i <= 2) then
        do
          xs = x + step
          call lag_eval(xs, n, alf, p, dp, p1)
          if (sign(1.0d0, pt) * sign(1.0d0, p) <= 0.0d0) then
            pt = p
            x = 0.5d0 * (x + xs)
            exit
          end if
          pt = p
          x = xs
        end do
      else
        fi = dble(i - 2)
        r1 = (1.0d0 + 2.55d0 * fi) / (1.9d0 * fi)
        r2 = 1.26d0 * fi * alf / (1.0d0 + 3.5d0 * fi)
        ratio = (r1 + r2) / (1.0d0 + 0.3d0 * alf)
        x = x + ratio * (x - xprev2)
      end if
      ! the weight uses p_n' and p_{n-1} of the last Newton evaluation
      do it = 1, 10
        call lag_eval(x, n, alf, p, dp, p1)
        d = p / dp
        x = x - d
        if (abs(d / x) <= 1.0d-12) exit
      end do
      xprev2 = xprev
      xprev = x
      y(i) = x
      w(i) = cc / dp / p1
      csx = csx + x
      csa = csa + w(i)
    end do
  end subroutine

  ! Points r, sqrt-weights, DVR kinetic matrix xk of one Morse-like radial
  ! coordinate (lagpt D:1290-1376 with zmorse=T, keints D:1495-1535 and k1k2
  ! D:1603-1632).  r = re + ln(A/y)/beta.
  subroutine radial_dvr(npnt, re, beta, a, iu, ur, r, xk, csx, csa, tsx, tsa, ierr)
    integer, intent(in) :: npnt, iu
    real(c_double), intent(in) :: re, beta, a, ur
    real(c_double), intent(out) :: r(npnt), xk(npnt, npnt), csx, csa, tsx, tsa
    integer, intent(out) :: ierr
    real(c_double) :: alf, cc, fke, y(npnt), wt(npnt), fact(0:npnt), binm(npnt + 1)
    real(c_double) :: dnorm(0:npnt - 1), bass(0:npnt - 1, npnt), t(0:npnt - 1, npnt)
    real(c_double) :: hbl(0:npnt - 1, 0:npnt - 1), cnt, amx, en, s
    integer :: i, m, mp, k, kp, nmax
    ierr = 0
    nmax = npnt - 1
    alf = dble(iu)
    ! normalisation of the Laguerre functions: dnorm(m)^2 = binom * m! * (n-m-1)!
    fact(0) = 1.0d0
    do i = 1, npnt
      fact(i) = dble(i) * fact(i - 1)
    end do
    cnt = dble(npnt) + alf
    cc = fact(npnt - 1) / cnt
    binm(npnt + 1) = 1.0d0
    do i = 1, npnt
      binm(npnt + 1 - i) = binm(npnt + 2 - i) * cnt / dble(i)
      cnt = cnt - 1.0d0
    end do
    do m = 0, nmax
      dnorm(m) = sqrt(binm(m + 1) * fact(m) * fact(npnt - m))
    end do
    call gauss_laguerre(npnt, alf, cc, y, wt, csx, csa, tsx)
    tsa = 1.0d0 / (dnorm(0) * dnorm(0))
    if (abs((csx - tsx) / tsx) > 1.0d-8 .or. abs((csa - tsa) / tsa) > 1.0d-8) ierr = 20
    do i = 1, npnt
      r(i) = re + log(a / y(i)) / beta
      ! normalised generalised Laguerre functions at the points
      bass(0, i) = 1.0d0
      if (nmax >= 1) then
        amx = alf + 1.0d0 - y(i)
        bass(1, i) = amx
        en = 1.0d0
        do m = 2, nmax
          en = en + 1.0d0
          amx = amx + 2.0d0
          bass(m, i) = (amx * bass(m - 1, i) - (alf - 1.0d0 + en) * bass(m - 2, i)) / en
        end do
      end if
      do m = 0, nmax
        bass(m, i) = bass(m, i) * dnorm(m)
      end do
      wt(i) = sqrt(wt(i))
    end do
    ! analytic kinetic energy in the FBR: diagonal and the n+2 band
    fke = beta * beta / (8.0d0 * ur)
    hbl = 0.0d0
    do m = 0, nmax
      hbl(m, m) = fke * dble(2 * m * (m + 1 + iu) + iu + 1)
      if (m + 2 <= nmax) then
        hbl(m, m + 2) = -fke * sqrt(dble((iu + m + 2) * (iu + m + 1)) * dble((m + 2) * (m + 1)))
        hbl(m + 2, m) = hbl(m, m + 2)
      end if
    end do
    ! DVR transformation T(m,k) = bass(m,k) w_k and xk = T^T hbl T
    do k = 1, npnt
      do m = 0, nmax
        t(m, k) = bass(m, k) * wt(k)
      end do
    end do
    do k = 1, npnt
      do kp = 1, k
        s = 0.0d0
        do m = 0, nmax
          do mp = 0, nmax
            s = s + hbl(m, mp) * t(m, k) * t(mp, kp)
          end do
        end do
        xk(k, kp) = s
        xk(kp, k) = s
      end do
    end do
  end subroutine

  ! ----------------------------------------------------------------- angular
  ! Monic Jacobi (alf = bta = nu) recurrence; returns p_n, p_n', p_{n-1}.
  subroutine jac_eval(x, n, nu, pn, dpn, pn1)
    real(c_double), intent(in) :: x, nu
    integer, intent(in) :: n
    real(c_double), intent(out) :: pn, dpn, pn1
    real(c_double) :: p0, p1, d0, d1, q, dq, cj, xj, s
    integer :: j
    p0 = 1.0d0; p1 = x
    d0 = 0.0d0; d1 = 1.0d0
    s = 2.0d0 * nu
    do j = 2, n
      xj = dble(j)
      cj = 4.0d0 * (xj - 1.0d0) * (nu + xj - 1.0d0) * (nu + xj - 1.0d0) * (s + xj - 1.0d0) / &
           ((s + 2.0d0 * xj - 1.0d0) * (s + 2.0d0 * xj - 2.0d0) * (s + 2.0d0 * xj - 2.0d0) * &
            (s + 2.0d0 * xj - 3.0d0))
      q = x * p1 - cj * p0
      dq = x * d1 + p1 - cj * d0
      p0 = p1; p1 = q; d0 = d1; d1 = dq
    end do
    pn = p1; dpn = d1; pn1 = p0
  end subroutine

  ! Positive-half Gauss-Jacobi (alf = bta = nu) points, from x = 1 downwards:
  ! roots bracketed by a sign scan (step = the single-precision 1e-4 of
  ! D:1685) and polished by Newton; weights cc / (p_n' p_{n-1}) with
  ! cc = (exact sum of weights) * prod c_j.
  subroutine gauss_jacobi_half(n, n2, nu, tsw, x, w, csw, ierr)
    integer, intent(in) :: n, n2
    real(c_double), intent(in) :: nu, tsw
    real(c_double), intent(out) :: x(n2), w(n2), csw
    integer, intent(out) :: ierr
    real(c_double) :: cc, xj, s, cj, xx, pm, p, dp, p1, d, stepd
    real(c_double) :: xt(n2)
    integer :: j, i, found, it
    ierr = 0
    s = 2.0d0 * nu
    cc = tsw
    do j = 2, n
      xj = dble(j)
      cj = 4.0d0 * (xj - 1.0d0) * (nu + xj - 1.0d0) * (nu + xj - 1.0d0) * (s + xj - 1.0d0) / &
           ((s + 2.0d0 * xj - 1.0d0) * (s + 2.0d0 * xj - 2.0d0) * (s + 2.0d0 * xj - 2.0d0) * &
            (s + 2.0d0 * xj - 3.0d0))
      cc = cc * cj
    end do
    stepd = dble(real(1.0e-4, kind=4))
    found = 0
    pm = 1.0d0
    xx = 1.0d0
    do
      call jac_eval(xx, n, nu, p, dp, p1)
      if (pm * p < 0.0d0) then
        pm = -pm
        found = found + 1
        xt(found) = xx
      end if
      if (found == n2) exit
      xx = xx - stepd
      if (.not. (xx > -0.002d0)) then
        ierr = 30
        return
      end if
    end do
    csw = 0.0d0
    do i = 1, n2
      xx = xt(i)
      do it = 1, 10
        call jac_eval(xx, n, nu, p, dp, p1)
        d = p / dp
        xx = xx - d
        if (abs(d) <= 1.0d-12) exit
      end do
      x(i) = xx
      w(i) = cc / (dp * p1)
      csw = csw + w(i) + w(i)
    end do
    if (2 * n2 /= n) csw = csw - w(n2)
  end subroutine

  ! Full angular basis of one k block: points xalf, sqrt-weights walf,
  ! normalised associated Legendre pleg(0:maxleg, idvr) and the L matrix.
  subroutine angular_dvr(job, sz, kz, xalf, walf, pleg, xl, csw, tsw, ierr)
    type(dvrh_job_t), intent(in) :: job
    type(dvrh_sizes_t), intent(in) :: sz
    integer, intent(in) :: kz
    real(c_double), intent(out) :: xalf(sz%idvr), walf(sz%idvr), pleg(0:sz%maxleg, sz%idvr)
    real(c_double), intent(out) :: xl(sz%idvr, sz%idvr), csw, tsw
    integer, intent(out) :: ierr
    real(c_double) :: xh(sz%nalf2), wh(sz%nalf2), scale, pmm, pmmp1, pll, fct, pnorm(0:sz%maxleg + kz)
    real(c_double) :: p(0:sz%maxleg + kz), s
    integer :: i, ia, l, ll, j, jj, jdia, jstart, jj0, k, kp, ipar0, lincr, nang, nang2
    ierr = 0
    lincr = kz
    nang = job%nalf
    nang2 = sz%nalf2
    ! exact sum of the Gauss-Jacobi weights, 2^(2k+1) (k!)^2 / (2k+1)!
    tsw = 2.0d0**(kz + kz + 1) / dble(kz + 1)
    do ia = 1, kz
      tsw = tsw * dble(ia) / dble(kz + ia + 1)
    end do
    call gauss_jacobi_half(nang, nang2, dble(kz), tsw, xh, wh, csw, ierr)
    if (ierr /= 0) return
    if (abs((csw - tsw) / tsw) > 1.0d-8) then
      ierr = 31; return
    end if
    scale = dble(max(1, job%idia))
    if (job%idia == 2) then                    ! half range, doubled weights
      do i = 1, nang2
        xalf(i) = xh(i)
        walf(i) = sqrt(scale * wh(i))
      end do
    else                                       ! mirror to the full range
      do i = 1, nang2
        walf(i) = sqrt(scale * wh(i))
        xalf(nang + 1 - i) = xh(i)
        xalf(i) = -xh(i)
        walf(nang + 1 - i) = walf(i)
      end do
    end if
    ! normalisation (2j+1)/2 (j-k)!/(j+k)!
    jj = -1
    do j = kz, sz%maxleg + lincr
      fct = 1.0d0
      do i = j - kz + 1, j + kz
        fct = fct * dble(i)
      end do
      jj = jj + 1
      pnorm(jj) = sqrt((dble(j) + dble(j) + 1.0d0) / (fct + fct))
    end do
    do i = 1, sz%idvr
      pmm = 1.0d0
      fct = 1.0d0
      do j = 1, kz
        pmm = -pmm * fct
        fct = fct + 2.0d0
      end do
      p(0) = pmm
      pmmp1 = xalf(i) * dble(kz + kz + 1) * pmm
      p(1) = pmmp1
      ll = 1
      do l = 2 + kz, sz%maxleg + lincr
        pll = (xalf(i) * dble(l + l - 1) * pmmp1 - dble(l + kz - 1) * pmm) / dble(l - kz)
        pmm = pmmp1
        pmmp1 = pll
        ll = ll + 1
        p(ll) = pll
      end do
      do jj = 0, sz%maxleg
        pleg(jj, i) = p(jj) * pnorm(jj)
      end do
    end do
    ! L(a,a') = w_a w_a' sum_j j(j+1) P_j(a) P_j(a'), j of the basis parity
    ipar0 = 0
    if (job%idia == 2 .and. sz%ipar == 1) ipar0 = 1
    jdia = max(1, job%idia)
    jstart = kz
    jj0 = -jdia
    if (mod(jstart, jdia) /= ipar0) then
      jj0 = jj0 + 1
      jstart = jstart + 1
    end if
    do k = 1, sz%idvr
      do kp = k, sz%idvr
        s = 0.0d0
        jj = jj0
        do j = jstart, sz%maxleg + lincr, jdia
          jj = jj + jdia
          s = s + pleg(jj, k) * pleg(jj, kp) * dble((j + 1) * j)
        end do
        xl(kp, k) = s * walf(k) * walf(kp)
        xl(k, kp) = xl(kp, k)
      end do
    end do
  end subroutine

  ! All basis arrays of one k block.  Arrays are allocated by the caller:
  ! r1(npnt1) r2(npnt2) xk1(npnt1,npnt1) xk2(npnt2,npnt2) xalf(idvr)
  ! walf(idvr) pleg(0:maxleg,idvr) xlmatr(idvr,idvr).
  integer(c_int) function dvrh_setup_basis(job, sz, kz, r1, r2, xk1, xk2, xalf, walf, &
                                           pleg, xl, cst) bind(C, name="dvrh_setup_basis")
    type(dvrh_job_t), intent(in) :: job
    type(dvrh_sizes_t), intent(in) :: sz
    integer(c_int), value :: kz
    real(c_double), intent(out) :: r1(job%npnt1), r2(job%npnt2)
    real(c_double), intent(out) :: xk1(job%npnt1, job%npnt1), xk2(job%npnt2, job%npnt2)
    real(c_double), intent(out) :: xalf(sz%idvr), walf(sz%idvr), pleg(0:sz%maxleg, sz%idvr)
    real(c_double), intent(out) :: xl(sz%idvr, sz%idvr)
    type(dvrh_consts_t), intent(out) :: cst
    integer :: ierr
    call set_consts(job, cst)
    call radial_dvr(job%npnt1, job%re1, cst%beta1, cst%a1, cst%iu1, cst%ur1, r1, xk1, &
                    cst%csx1, cst%csa1, cst%tsx1, cst%tsa1, ierr)
    if (ierr /= 0) then
      dvrh_setup_basis = ierr; return
    end if
    call radial_dvr(job%npnt2, job%re2, cst%beta2, cst%a2, cst%iu2, cst%ur2, r2, xk2, &
                    cst%csx2, cst%csa2, cst%tsx2, cst%tsa2, ierr)
    if (ierr /= 0) then
      dvrh_setup_basis = ierr; return
    end if
    call angular_dvr(job, sz, kz, xalf, walf, pleg, xl, cst%cswalf, cst%tswalf, ierr)
    dvrh_setup_basis = ierr
  end function

  ! Set-up of the symmetrised-Radau driver (IDIA = -2, J = 0): the sizes of
  ! insize (npnt1 = npnt2, idvr = nalf, D:246-252, 333-345), the r1 radial
  ! DVR used for both bonds and the full-range angular DVR with its L
  ! matrix.  Returns 0, 10 for J > 0 or other options this build does not
  ! take for IDIA = -2, or the quadrature's error code.
  integer(c_int) function dvrh_setup_nf(job, sz, r, hr, xalf, walf, pleg, xl, cst) &
      bind(C, name="dvrh_setup_nf")
    type(dvrh_job_t), intent(in) :: job
    type(dvrh_sizes_t), intent(inout) :: sz
    real(c_double), intent(out) :: r(job%npnt2), hr(job%npnt2, job%npnt2)
    real(c_double), intent(out) :: xalf(job%nalf), walf(job%nalf), pleg(0:job%nalf - 1, job%nalf)
    real(c_double), intent(out) :: xl(job%nalf, job%nalf)
    type(dvrh_consts_t), intent(out) :: cst
    integer :: ierr, npair
    dvrh_setup_nf = 0
    if (job%idia /= -2 .or. job%jrot /= 0 .or. job%zperp /= 0 .or. job%ztwod /= 0 .or. &
        job%zmors1 == 0 .or. job%zmors2 == 0 .or. (job%ipar /= 0 .and. job%ipar /= 1)) then
      dvrh_setup_nf = 10; return
    end if
    sz%nalf2 = (job%nalf + 1) / 2
    sz%maxleg = job%nalf - 1
    sz%idvr = job%nalf
    sz%npnta = job%nalf
    sz%ipar = job%ipar
    sz%idia = -2; sz%jrot = 0; sz%kmin = 0; sz%zrot = 0; sz%zembed = 1
    sz%kd = 0; sz%ku = 0
    npair = job%npnt2 * (job%npnt2 + 1 - 2 * job%ipar) / 2
    sz%max2d = npair
    if (job%zall /= 0) then
      sz%max3d = npair * job%nalf
    else
      sz%max3d = min(job%max3d, npair * job%nalf)
    end if
    sz%neval = job%neval
    if (sz%neval <= 0) sz%neval = 10
    sz%neval = min(sz%max3d, sz%neval)
    sz%zvec = job%zvec
    call set_consts(job, cst)
    call radial_dvr(job%npnt2, job%re1, cst%beta1, cst%a1, cst%iu1, cst%ur1, r, hr, &
                    cst%csx1, cst%csa1, cst%tsx1, cst%tsa1, ierr)
    if (ierr /= 0) then
      dvrh_setup_nf = ierr; return
    end if
    call angular_dvr(job, sz, 0, xalf, walf, pleg, xl, cst%cswalf, cst%tswalf, ierr)
    dvrh_setup_nf = ierr
  end function

  ! (w1 + w2) and the rotational diagonal term of mkham1 at every (c, b)
  ! point, theta-first (D:2151-2173); index c*npntb + b.  wterm is zero for
  ! J = 0.
  integer(c_int) function dvrh_wsum(job, sz, cst, kz, r1, r2, wsum, wterm) &
      bind(C, name="dvrh_wsum")
    type(dvrh_job_t), intent(in) :: job
    type(dvrh_sizes_t), intent(in) :: sz
    type(dvrh_consts_t), intent(in) :: cst
    integer(c_int), value :: kz
    real(c_double), intent(in) :: r1(job%npnt1), r2(job%npnt2)
    real(c_double), intent(out) :: wsum(0:sz%npntb * sz%npntc - 1), wterm(0:sz%npntb * sz%npntc - 1)
    real(c_double) :: w1, w2, term
    integer :: c, b, i1, i2
    term = dble(sz%jrot * sz%jrot + sz%jrot - 2 * kz * kz)
    do c = 0, sz%npntc - 1
      do b = 0, sz%npntb - 1
        if (job%zr2r1 /= 0) then
          i1 = c + 1; i2 = b + 1
        else
          i1 = b + 1; i2 = c + 1
        end if
        w1 = 0.5d0 / (r1(i1) * r1(i1) * cst%ur1)
        w2 = 0.5d0 / (r2(i2) * r2(i2) * cst%ur2)
        wsum(c * sz%npntb + b) = w1 + w2
        wterm(c * sz%npntb + b) = 0.0d0
        if (sz%jrot > 0) then
          if (sz%zembed /= 0) then
            wterm(c * sz%npntb + b) = term * w2 * cst%ur2 / cst%urr2
          else
            wterm(c * sz%npntb + b) = term * w1 * cst%ur1 / cst%urr1
          end if
        end if
      end do
    end do
    dvrh_wsum = 0
  end function

end module dvrh_setup
