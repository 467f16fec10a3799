! dsyev with its optimal workspace, for the "fast" build of the reference
! program (oracle/Makefile target fast).  Test infrastructure only.
!
! The reference passes LWORK = 3 N to dsyev from diag and diag3d
! (SFIT_DVR3D/H2S/dvr3drjz.f90:2591-2593, 2639-2640; the same in
! source/dvr3drjz_segmented.f90:2650-2654), which makes dsytrd and dorgtr
! fall back to their unblocked forms (source/dsyev.f:174-188, 6398-6424).
! The fast build calls this routine instead: it asks dsyev for the optimal
! workspace (LWORK = -1), allocates it and calls dsyev again.  The
! reference's own work array is not used.
      subroutine dsyev_optlw(jobz, uplo, n, a, lda, w, work, lwork, info)
      implicit none
      character, intent(in) :: jobz, uplo
      integer, intent(in) :: n, lda, lwork
      integer, intent(out) :: info
      double precision, intent(inout) :: a(lda, *)
      double precision, intent(out) :: w(*)
      double precision, intent(inout) :: work(*)
      double precision :: q(1)
      double precision, allocatable :: wk(:)
      integer :: lw
      call dsyev(jobz, uplo, n, a, lda, w, q, -1, info)
      lw = max(int(q(1)), 3 * n, lwork, 1)
      allocate(wk(lw))
      call dsyev(jobz, uplo, n, a, lda, w, wk, lw, info)
      deallocate(wk)
      end subroutine dsyev_optlw
