! C entry points over the reference's regularizer module, for the tests
! only.  Built with the reference's own src/fortlib/param.f90 and
! src/fortlib/image.f90 (compiled from their location under the reference
! tree by oracle/Makefile, target `ref`) into oracle/_ref/libsmili_ref.so;
! nothing of the reference is copied here.  Serial build (no OpenMP): the
! sums run in pixel order.
module ref_shim
  use iso_c_binding
  use param, only: dp
  use image, only: calc_cost_reg, init_l1reg, init_tvreg, init_tsvreg, init_klreg, &
                   init_gsreg, init_tfdreg, calc_tfdreg, calc_cenreg
  implicit none
contains

! image.calc_cost_reg (image.f90:12-200).  Every weight vector has N1d
! entries (the caller passes any values for disabled terms).  costs(1:11) =
! l1, tv, tsv, kl, gs, tfd, cen, sm, out_maj, out_min, out_phi.
subroutine ref_calc_cost_reg(N1d, Nx, Ny, I1d, xidx, yidx, Nxref, Nyref, &
    l1_l, l1_wgt, tv_l, tv_wgt, tsv_l, tsv_wgt, kl_l, kl_wgt, gs_l, gs_wgt, &
    tfd_l, tfd_tgtfd, cen_l, cen_alpha, sm_l, sm_maj, sm_min, sm_phi, &
    costs, cost, gradcost) bind(C, name="ref_calc_cost_reg")
  integer(c_int), value :: N1d, Nx, Ny
  real(c_double), intent(in) :: I1d(N1d)
  integer(c_int), intent(in) :: xidx(N1d), yidx(N1d)
  real(c_double), value :: Nxref, Nyref, l1_l, tv_l, tsv_l, kl_l, gs_l
  real(c_double), intent(in) :: l1_wgt(N1d), tv_wgt(N1d), tsv_wgt(N1d), kl_wgt(N1d), gs_wgt(N1d)
  real(c_double), value :: tfd_l, tfd_tgtfd, cen_l, cen_alpha, sm_l, sm_maj, sm_min, sm_phi
  real(c_double), intent(out) :: costs(11), cost, gradcost(N1d)
  call calc_cost_reg(I1d, xidx, yidx, Nxref, Nyref, Nx, Ny, &
      l1_l, l1_wgt, N1d, tv_l, tv_wgt, N1d, tsv_l, tsv_wgt, N1d, &
      kl_l, kl_wgt, N1d, gs_l, gs_wgt, N1d, &
      tfd_l, tfd_tgtfd, cen_l, cen_alpha, sm_l, sm_maj, sm_min, sm_phi, &
      costs(1), costs(2), costs(3), costs(4), costs(5), costs(6), costs(7), costs(8), &
      costs(9), costs(10), costs(11), cost, gradcost, N1d)
end subroutine

subroutine ref_init_l1reg(N1d, prior, wgt) bind(C, name="ref_init_l1reg")
  integer(c_int), value :: N1d
  real(c_double), intent(in) :: prior(N1d)
  real(c_double), intent(out) :: wgt(N1d)
  call init_l1reg(prior, wgt, N1d)
end subroutine

subroutine ref_init_tvreg(N1d, Nx, Ny, xidx, yidx, isflat, prior, wgt) bind(C, name="ref_init_tvreg")
  integer(c_int), value :: N1d, Nx, Ny, isflat
  integer(c_int), intent(in) :: xidx(N1d), yidx(N1d)
  real(c_double), intent(in) :: prior(N1d)
  real(c_double), intent(out) :: wgt(N1d)
  call init_tvreg(xidx, yidx, Nx, Ny, isflat /= 0, prior, wgt, N1d)
end subroutine

subroutine ref_init_tsvreg(N1d, Nx, Ny, xidx, yidx, isflat, prior, wgt) bind(C, name="ref_init_tsvreg")
  integer(c_int), value :: N1d, Nx, Ny, isflat
  integer(c_int), intent(in) :: xidx(N1d), yidx(N1d)
  real(c_double), intent(in) :: prior(N1d)
  real(c_double), intent(out) :: wgt(N1d)
  call init_tsvreg(xidx, yidx, Nx, Ny, isflat /= 0, prior, wgt, N1d)
end subroutine

subroutine ref_init_klreg(N1d, l_in, prior, l_out, wgt) bind(C, name="ref_init_klreg")
  integer(c_int), value :: N1d
  real(c_double), value :: l_in
  real(c_double), intent(in) :: prior(N1d)
  real(c_double), intent(out) :: l_out, wgt(N1d)
  call init_klreg(l_in, prior, l_out, wgt, N1d)
end subroutine

subroutine ref_init_gsreg(N1d, prior, wgt) bind(C, name="ref_init_gsreg")
  integer(c_int), value :: N1d
  real(c_double), intent(in) :: prior(N1d)
  real(c_double), intent(out) :: wgt(N1d)
  call init_gsreg(prior, wgt, N1d)
end subroutine

subroutine ref_init_tfdreg(l_in, tgtfd, tgter, l_out) bind(C, name="ref_init_tfdreg")
  real(c_double), value :: l_in, tgtfd, tgter
  real(c_double), intent(out) :: l_out
  call init_tfdreg(l_in, tgtfd, tgter, l_out)
end subroutine

subroutine ref_calc_cenreg(N1d, I1d, xidx, yidx, Nxref, Nyref, alpha, cost, grad) &
    bind(C, name="ref_calc_cenreg")
  integer(c_int), value :: N1d
  real(c_double), intent(in) :: I1d(N1d)
  integer(c_int), intent(in) :: xidx(N1d), yidx(N1d)
  real(c_double), value :: Nxref, Nyref, alpha
  real(c_double), intent(out) :: cost, grad(N1d)
  call calc_cenreg(I1d, xidx, yidx, Nxref, Nyref, alpha, cost, grad, N1d)
end subroutine

end module
