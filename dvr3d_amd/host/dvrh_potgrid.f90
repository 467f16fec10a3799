! The linked POTV on the grid of one k block, theta first: v(a, b, c) with a
! the angle, b and c the radial coordinates in the order of zr2r1 -- the loop
! of mkham1 (source/dvr3drjz_segmented.f90:2178) that the drop-in executable
! runs before handing the grid to dvr_set_potential.  COMMON /MASS/ is set
! from the job and the basis constants first, as setcon does for POTV.
integer(c_int) function dvrh_potential_grid(job, sz, cst, r1, r2, xalf, v) &
    bind(C, name="dvrh_potential_grid")
  use iso_c_binding
  use dvrh_setup
  implicit none
  type(dvrh_job_t), intent(in) :: job
  type(dvrh_sizes_t), intent(in) :: sz
  type(dvrh_consts_t), intent(in) :: cst
  real(c_double), intent(in) :: r1(job%npnt1), r2(job%npnt2), xalf(sz%idvr)
  real(c_double), intent(out) :: v(sz%npnta, sz%npntb, sz%npntc)
  double precision :: xmass(3), g1, g2, xmassr(3), vv
  common /mass/ xmass, g1, g2, xmassr
  integer :: a, b, c
  xmass = job%xmass; xmassr = job%xmassr
  g1 = cst%g1; g2 = cst%g2
  do c = 1, sz%npntc
    do b = 1, sz%npntb
      do a = 1, sz%npnta
        if (job%zr2r1 /= 0) then
          call potv(vv, r1(c), r2(b), xalf(a))
        else
          call potv(vv, r1(b), r2(c), xalf(a))
        end if
        v(a, b, c) = vv
      end do
    end do
  end do
  dvrh_potential_grid = 0
end function
