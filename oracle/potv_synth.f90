! Synthetic potential-energy surfaces for the DVR3DRJZ oracle runs.
!
! Test infrastructure only: this file is linked into the reference
! executable built by oracle/Makefile (oracle/_ref/), which serves as the
! FP64 parity oracle and CPU baseline.  It is the user "plugin" that the
! reference expects at link time:
!     SUBROUTINE POTV(V,R1,R2,XCOS)   with   COMMON /MASS/ XMASS(3),G1,G2,XMASSR(3)
! (docum.tex:31-45; called from source/dvr3drjz_segmented.f90:2178).
!
! Parameters come from the text file named by the environment variable
! DVR_PES (default "pes.par" in the working directory), read once:
!     line 1: kind            (1 = symmetric AB2 Morse + bend + cross term)
!     line 2: De a re         (Eh, 1/bohr, bohr)
!     line 3: kb thetae_deg c (Eh, degrees, Eh)
!
! kind 2 (ABC in Jacobi coordinates, IDIA=1; line 2: De a re1, line 3:
!   kb re2 noise):  V = De[(1-exp(-a(r1-re1)))^2 + (1-exp(-a(d-re2)))^2]
!   + kb (cos t - 1)^2 with d = sqrt(yy^2 + r2^2 + 2 yy r2 cos t),
!   yy = r1 (1 - g1): the atom 1 - atom 2 distance.  Only noise = 0 is
!   available here (the hashed noise is restated in the numpy oracle).
! kind 1 (AB2, Radau or Jacobi coordinates, r1/r2 symmetric):
!   V = De[(1-exp(-a(r1-re)))^2 + (1-exp(-a(r2-re)))^2]
!       + kb (cos t - cos te)^2 + c (r1-re)(r2-re)
! The two Morse terms are added as m1+m2 and the cross term as a product,
! so V(r1,r2,x) == V(r2,r1,x) holds bit for bit.
      subroutine potv(v, r1, r2, xcos)
      implicit none
      double precision, intent(out) :: v
      double precision, intent(in) :: r1, r2, xcos
      double precision :: xmass(3), g1, g2, xmassr(3)
      common /mass/ xmass, g1, g2, xmassr
      integer, save :: kind = 0
      double precision, save :: de, a, re, kb, te, c, cte
      logical, save :: loaded = .false.
      character(len=512) :: fname
      integer :: n, ios
      double precision :: m1, m2, yy, dab
      if (.not. loaded) then
         call get_environment_variable('DVR_PES', fname, n, ios)
         if (ios /= 0 .or. n == 0) fname = 'pes.par'
         open(unit=91, file=trim(fname), status='old', action='read', iostat=ios)
         if (ios /= 0) then
            write(6,*) 'potv: cannot open PES parameter file ', trim(fname)
            stop
         endif
         read(91,*) kind
         read(91,*) de, a, re
         read(91,*) kb, te, c
         close(91)
         cte = cos(te * (4.0d0*atan(1.0d0)) / 180.0d0)
         loaded = .true.
      endif
      select case (kind)
      case (1)
         m1 = (1.0d0 - exp(-a*(r1-re)))**2
         m2 = (1.0d0 - exp(-a*(r2-re)))**2
         v = de*(m1+m2) + kb*(xcos-cte)**2 + c*((r1-re)*(r2-re))
      case (2)
         if (c /= 0.0d0) then
            write(6,*) 'potv: kind 2 noise is not available in this build'
            stop
         endif
         yy = r1 * (1.0d0 - g1)
         dab = sqrt(yy*yy + r2*r2 + 2.0d0*yy*r2*xcos)
         m1 = (1.0d0 - exp(-a*(r1-re)))**2
         m2 = (1.0d0 - exp(-a*(dab-te)))**2
         v = de*(m1+m2) + kb*(xcos-1.0d0)**2
      case default
         write(6,*) 'potv: unknown PES kind', kind
         stop
      end select
      end subroutine potv
