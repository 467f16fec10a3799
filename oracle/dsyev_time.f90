! Wall time of the fast reference build's eigensolver call on one matrix
! (oracle/Makefile, linked with dsyev_optlw.f90 and MKL sequential like
! dvr3drjz_fast).  Used by bench.py to scale the reference's final dsyev
! (diag3d, SFIT_DVR3D/H2S/dvr3drjz.f90:2640) to dimensions it cannot be run
! at in a benchmark.  Test infrastructure only.
!
!   dsyev_time N   ->  prints "N seconds"
!
! The matrix is a deterministic symmetric one with a spread diagonal, like a
! contracted Hamiltonian: a(i,j) = 0.01 cos(i + 2 j) + 0.01 cos(2 i + j)
! off the diagonal, i / N on it.
      program dsyev_time
      implicit none
      integer :: n, i, j, info, c0, c1, rate
      character(len=32) :: arg
      double precision, allocatable :: a(:, :), w(:), wk(:)
      call get_command_argument(1, arg)
      read (arg, *) n
      allocate(a(n, n), w(n), wk(3 * n))
      do j = 1, n
        do i = 1, n
          a(i, j) = 0.01d0 * cos(dble(i + 2 * j)) + 0.01d0 * cos(dble(2 * i + j))
        end do
        a(j, j) = dble(j) / dble(n)
      end do
      call system_clock(c0, rate)
      call dsyev_optlw('V', 'L', n, a, n, w, wk, 3 * n, info)
      call system_clock(c1)
      write (*, '(i8, f14.4, i4)') n, dble(c1 - c0) / dble(rate), info
      end program dsyev_time
