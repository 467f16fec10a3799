! ISO_C_BINDING interface of libdvr3d_amd (include/dvr3d_amd.h).
module dvr3d_amd_api
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: DVR_OK = 0, DVR_ERR_ARG = 1, DVR_ERR_HIP = 2
  integer(c_int), parameter :: DVR_ERR_CORE_2D = 3, DVR_ERR_CORE_3D = 4, DVR_ERR_NOCONV = 5
  integer(c_int), parameter :: DVR_SOLVER_DENSE = 0, DVR_SOLVER_DAVIDSON = 1
  integer(c_int), parameter :: DVR_NSTAGE = 7

  type, bind(C) :: dvr_dims_t
    integer(c_int) :: npnta, npntb, npntc
    integer(c_int) :: max2d, max3d, neval
    integer(c_int) :: zcut, zall, zr2r1
    integer(c_int) :: idvr, npnt1, npnt2
    integer(c_int) :: nvec, solver
  end type

  type, bind(C) :: dvr_nf_dims_t
    integer(c_int) :: npnt, nalf, ipar
    integer(c_int) :: max3d, neval
    integer(c_int) :: zcut, nvec
    integer(c_int) :: jrot, kz, kmin, zrot, zlin
  end type

  type, bind(C) :: dvr_rot_dims_t
    integer(c_int) :: nblk, kfirst, jrot, idia, ipar
    integer(c_int) :: npnt1, npnt2, zembed, zquad2, nlim
    integer(c_int) :: nsets
    integer(c_int) :: neval(2)
    integer(c_int) :: nvec
  end type

  type, bind(C) :: dvr_rot_block_t
    integer(c_int) :: maxleg, idvr, lincr, mvib
  end type

  type, bind(C) :: dvr_dip_dims_t
    integer(c_int) :: nrad, ipot, ne1, ne2, nblk1, nblk2, npair
  end type

  type, bind(C) :: dvr_dip_block_t
    integer(c_int) :: nin
    type(c_ptr) :: vec, ang
  end type

  type, bind(C) :: dvr_dip_pair_t
    integer(c_int) :: b1, b2, comp
    real(c_double) :: factor
  end type

  type, bind(C) :: dvr_xp_dims_t
    integer(c_int) :: npts, npot, ne, nblk, nprop
  end type

  type, bind(C) :: dvr_xp_block_t
    integer(c_int) :: nin
    integer(c_int64_t) :: ldv
    type(c_ptr) :: vec, ang
  end type

  interface
    integer(c_int) function dvr_xpect3(dims, blocks, prop, out) bind(C, name="dvr_xpect3")
      import :: c_int, c_double, dvr_xp_dims_t, dvr_xp_block_t
      type(dvr_xp_dims_t), intent(in) :: dims
      type(dvr_xp_block_t), intent(in) :: blocks(*)
      real(c_double), intent(in) :: prop(*)
      real(c_double), intent(inout) :: out(*)
    end function
    integer(c_int) function dvr_dipole3(dims, ket, bra, dipol, pairs, tz, tx, stats) &
        bind(C, name="dvr_dipole3")
      import :: c_int, c_double, c_ptr, dvr_dip_dims_t, dvr_dip_block_t, dvr_dip_pair_t
      type(dvr_dip_dims_t), intent(in) :: dims
      type(dvr_dip_block_t), intent(in) :: ket(*), bra(*)
      real(c_double), intent(in) :: dipol(*)
      type(dvr_dip_pair_t), intent(in) :: pairs(*)
      real(c_double), intent(inout) :: tz(*), tx(*)
      type(c_ptr), value :: stats
    end function
    integer(c_int) function dvr_rotlev3(dims, blocks, pleg, dvrvec, evib, rm2, jst_nang, &
                                        eval, vec, dvec) bind(C, name="dvr_rotlev3")
      import :: c_int, c_double, dvr_rot_dims_t, dvr_rot_block_t
      type(dvr_rot_dims_t), intent(in) :: dims
      type(dvr_rot_block_t), intent(in) :: blocks(*)
      real(c_double), intent(in) :: pleg(*), dvrvec(*), evib(*), rm2(*)
      integer(c_int), intent(out) :: jst_nang(*)
      real(c_double), intent(out) :: eval(*), vec(*), dvec(*)
    end function
    integer(c_int) function dvr_nfmain(dims, hr, htheta, r, xalf, ur1, urr1, vgrid, emax2, &
                                       iv2, nham3, iang, eval, wvfunc) bind(C, name="dvr_nfmain")
      import :: c_int, c_double, dvr_nf_dims_t
      type(dvr_nf_dims_t), intent(in) :: dims
      real(c_double), intent(in) :: hr(*), htheta(*), r(*), xalf(*), vgrid(*)
      real(c_double), value :: ur1, urr1, emax2
      integer(c_int), intent(out) :: iv2(*), nham3, iang
      real(c_double), intent(out) :: eval(*), wvfunc(*)
    end function
    integer(c_int) function dvr_ctx_create(dims, device, ctx) bind(C, name="dvr_ctx_create")
      import :: c_int, c_ptr, dvr_dims_t
      type(dvr_dims_t), intent(in) :: dims
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function
    subroutine dvr_ctx_destroy(ctx) bind(C, name="dvr_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function dvr_last_error(ctx) bind(C, name="dvr_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function dvr_set_basis(ctx, xlmatr, xkb, xkc, wsum, wterm, emax1, emax2) &
        bind(C, name="dvr_set_basis")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: xlmatr(*), xkb(*), xkc(*), wsum(*), wterm(*)
      real(c_double), value :: emax1, emax2
    end function
    integer(c_int) function dvr_set_potential(ctx, v) bind(C, name="dvr_set_potential")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: v(*)
    end function
    integer(c_int) function dvr_run(ctx) bind(C, name="dvr_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function dvr_run_blocks(ctxs, n) bind(C, name="dvr_run_blocks")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: ctxs(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function dvr_get_sizes(ctx, iv1, iv2, nham3) bind(C, name="dvr_get_sizes")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: iv1(*), iv2(*), nham3
    end function
    integer(c_int) function dvr_get_eigenvalues(ctx, eval, meval) bind(C, name="dvr_get_eigenvalues")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: eval(*)
      integer(c_int), intent(out) :: meval
    end function
    integer(c_int) function dvr_get_wavefunctions(ctx, psi) bind(C, name="dvr_get_wavefunctions")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: psi(*)
    end function
    integer(c_int) function dvr_get_stage_times(ctx, ms) bind(C, name="dvr_get_stage_times")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(*)
    end function
    integer(c_int) function dvr_diag(ham, maxham, nham, eig) bind(C, name="dvr_diag")
      import :: c_int, c_double
      real(c_double), intent(inout) :: ham(*)
      integer(c_int), value :: maxham, nham
      real(c_double), intent(out) :: eig(*)
    end function
  end interface

contains

  ! Message of the last failed call on ctx, as a Fortran string.
  function dvr_error_text(ctx) result(s)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: s
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: i
    s = ' '
    cp = dvr_last_error(ctx)
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [256])
    do i = 1, 256
      if (p(i) == c_null_char) exit
      s(i:i) = p(i)
    end do
  end function

end module dvr3d_amd_api
