! ISO_C_BINDING interface to libafivo_amd (include/afivo_amd.h).
!
! Dimension-independent; used by the drop-in modules m_a2_mg and m_a3_mg.
module m_afg_capi
  use iso_c_binding
  implicit none
  public

  integer(c_int), parameter :: AFG_OP_LPL = 0, AFG_OP_AUTO = 1, AFG_OP_LSF = 2
  integer(c_int), parameter :: AFG_RSTR_LPL = 0, AFG_RSTR_AUTO = 1, AFG_RSTR_LSF = 2
  integer(c_int), parameter :: AFG_BC_DIRICHLET = -1, AFG_BC_NEUMANN = -2
  integer(c_int), parameter :: AFG_COORD_XYZ = 0, AFG_COORD_CYL = 1
  integer(c_int), parameter :: AFG_BOX_LPL = 0, AFG_BOX_LPLD = 1, AFG_BOX_LSF = 2, AFG_BOX_CLPL = 3, &
       AFG_BOX_CLPLD = 4

  !> afg_mg_opts of include/afivo_amd.h
  type, bind(c) :: afg_mg_opts
     integer(c_int32_t) :: i_phi = -1, i_rhs = -1, i_tmp = -1, i_eps = -1
     integer(c_int32_t) :: n_cycle_down = -1, n_cycle_up = -1, n_cycle_base = -1
     integer(c_int32_t) :: op = 0, initialized = 0
     integer(c_int32_t) :: i_lsf = -1, i_bval = -1, rstr = 0
     integer(c_int32_t) :: reserved(4) = 0
  end type afg_mg_opts

  interface
     integer(c_int) function afg_ctx_create(dim, n_cell, n_var, ctx) bind(c, name="afg_ctx_create")
       import :: c_int, c_ptr
       integer(c_int), value :: dim, n_cell, n_var
       type(c_ptr), intent(out) :: ctx
     end function afg_ctx_create

     integer(c_int) function afg_ctx_destroy(ctx) bind(c, name="afg_ctx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function afg_ctx_destroy

     type(c_ptr) function afg_last_error(ctx) bind(c, name="afg_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function afg_last_error

     integer(c_int) function afg_set_tree(ctx, n_boxes, records, dr_base) bind(c, name="afg_set_tree")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_boxes
       integer(c_int32_t), intent(in) :: records(*)
       real(c_double), value :: dr_base
     end function afg_set_tree

     integer(c_int) function afg_set_coord(ctx, coord, r_base) bind(c, name="afg_set_coord")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: coord
       real(c_double), value :: r_base
     end function afg_set_coord

     integer(c_int) function afg_update_tree(ctx, n_boxes, records, dr_base, tree_of, prolong_order, &
          n_vars, vars) bind(c, name="afg_update_tree")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_boxes
       integer(c_int32_t), intent(in) :: records(*)
       real(c_double), value :: dr_base
       type(c_ptr), value :: tree_of
       integer(c_int), value :: prolong_order, n_vars
       integer(c_int32_t), intent(in) :: vars(*)
     end function afg_update_tree

     integer(c_int) function afg_set_bc(ctx, id, nb, bc_type, b) bind(c, name="afg_set_bc")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: id, nb, bc_type
       type(c_ptr), value :: b
     end function afg_set_bc

     ! per-box operators on the host (box arrays cc(0:nc+1, ...) of one variable)
     integer(c_int) function afg_box_op(dim, nc, dr, kind, ir0, phi, eps, lsf, bval, out) &
          bind(c, name="afg_box_op")
       import :: c_int, c_double
       integer(c_int), value :: dim, nc, kind, ir0
       real(c_double), value :: dr
       real(c_double), intent(in) :: phi(*), eps(*), lsf(*), bval(*)
       real(c_double), intent(inout) :: out(*)
     end function afg_box_op

     integer(c_int) function afg_box_gsrb(dim, nc, dr, kind, ir0, redblack, phi, rhs, eps, lsf, bval) &
          bind(c, name="afg_box_gsrb")
       import :: c_int, c_double
       integer(c_int), value :: dim, nc, kind, ir0, redblack
       real(c_double), value :: dr
       real(c_double), intent(inout) :: phi(*)
       real(c_double), intent(in) :: rhs(*), eps(*), lsf(*), bval(*)
     end function afg_box_gsrb

     integer(c_int) function afg_box_corr(dim, nc, pnc, kind, offset, p_corr, p_eps, p_lsf, c_phi, c_lsf) &
          bind(c, name="afg_box_corr")
       import :: c_int, c_int32_t, c_double
       integer(c_int), value :: dim, nc, pnc, kind
       integer(c_int32_t), intent(in) :: offset(*)
       real(c_double), intent(in) :: p_corr(*), p_eps(*), p_lsf(*), c_lsf(*)
       real(c_double), intent(inout) :: c_phi(*)
     end function afg_box_corr

     integer(c_int) function afg_box_rstr(dim, nc, pnc, kind, offset, c_var, c_lsf, p_lsf, p_rmin, c_dr, p_var) &
          bind(c, name="afg_box_rstr")
       import :: c_int, c_int32_t, c_double
       integer(c_int), value :: dim, nc, pnc, kind
       integer(c_int32_t), intent(in) :: offset(*)
       real(c_double), intent(in) :: c_var(*), c_lsf(*), p_lsf(*)
       real(c_double), value :: p_rmin, c_dr
       real(c_double), intent(inout) :: p_var(*)
     end function afg_box_rstr

     integer(c_int) function afg_box_sides_rb(dim, nc, pnc, nb, offset, p_var, c_var) &
          bind(c, name="afg_box_sides_rb")
       import :: c_int, c_int32_t, c_double
       integer(c_int), value :: dim, nc, pnc, nb
       integer(c_int32_t), intent(in) :: offset(*)
       real(c_double), intent(in) :: p_var(*)
       real(c_double), intent(inout) :: c_var(*)
     end function afg_box_sides_rb

     integer(c_int) function afg_upload_cc(ctx, iv, n_ids, ids, packed) bind(c, name="afg_upload_cc")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: iv, n_ids
       integer(c_int32_t), intent(in) :: ids(*)
       real(c_double), intent(in) :: packed(*)
     end function afg_upload_cc

     integer(c_int) function afg_download_cc(ctx, iv, n_ids, ids, packed) bind(c, name="afg_download_cc")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: iv, n_ids
       integer(c_int32_t), intent(in) :: ids(*)
       real(c_double), intent(out) :: packed(*)
     end function afg_download_cc

     integer(c_int) function afg_upload_boxes(ctx, iv, n_ids, ids, boxes) bind(c, name="afg_upload_boxes")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: iv, n_ids
       integer(c_int32_t), intent(in) :: ids(*)
       type(c_ptr), intent(in) :: boxes(*)
     end function afg_upload_boxes

     integer(c_int) function afg_download_boxes(ctx, iv, n_ids, ids, boxes) bind(c, name="afg_download_boxes")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: iv, n_ids
       integer(c_int32_t), intent(in) :: ids(*)
       type(c_ptr), intent(in) :: boxes(*)
     end function afg_download_boxes

     integer(c_int) function afg_mg_init(opts) bind(c, name="afg_mg_init")
       import :: c_int, afg_mg_opts
       type(afg_mg_opts), intent(inout) :: opts
     end function afg_mg_init

     integer(c_int) function afg_fas_fmg(ctx, opts, set_residual, have_guess) bind(c, name="afg_fas_fmg")
       import :: c_int, c_ptr, afg_mg_opts
       type(c_ptr), value :: ctx
       type(afg_mg_opts), intent(in) :: opts
       integer(c_int), value :: set_residual, have_guess
     end function afg_fas_fmg

     integer(c_int) function afg_fas_vcycle(ctx, opts, highest_lvl, set_residual) bind(c, name="afg_fas_vcycle")
       import :: c_int, c_ptr, afg_mg_opts
       type(c_ptr), value :: ctx
       type(afg_mg_opts), intent(in) :: opts
       integer(c_int), value :: highest_lvl, set_residual
     end function afg_fas_vcycle

     integer(c_int) function afg_mixed_solve(ctx, opts, have_guess, tol, max_iter, cycles, iterations, &
          rel_residual) bind(c, name="afg_mixed_solve")
       import :: c_int, c_ptr, c_double, afg_mg_opts
       type(c_ptr), value :: ctx
       type(afg_mg_opts), intent(in) :: opts
       integer(c_int), value :: have_guess, max_iter, cycles
       real(c_double), value :: tol
       integer(c_int), intent(out) :: iterations
       real(c_double), intent(out) :: rel_residual
     end function afg_mixed_solve

     integer(c_int) function afg_autotune(ctx, opts, fp32, reps, ms) bind(c, name="afg_autotune")
       import :: c_int, c_ptr, afg_mg_opts
       type(c_ptr), value :: ctx
       type(afg_mg_opts), intent(in) :: opts
       integer(c_int), value :: fp32, reps
       type(c_ptr), value :: ms
     end function afg_autotune

     integer(c_int) function afg_tree_max_cc(ctx, iv, include_parents, out) bind(c, name="afg_tree_max_cc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: iv, include_parents
       real(c_double), intent(out) :: out
     end function afg_tree_max_cc

     integer(c_int) function afg_tree_min_cc(ctx, iv, include_parents, out) bind(c, name="afg_tree_min_cc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: iv, include_parents
       real(c_double), intent(out) :: out
     end function afg_tree_min_cc

     integer(c_size_t) function afg_strlen(s) bind(c, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function afg_strlen
  end interface

contains

  !> Stop with the library's message when a call failed.
  subroutine afg_check(rc, ctx, where)
    integer(c_int), intent(in)   :: rc
    type(c_ptr), intent(in)      :: ctx
    character(len=*), intent(in) :: where
    type(c_ptr)                  :: p
    character(kind=c_char), pointer :: chars(:)
    character(len=:), allocatable :: msg
    integer :: n, i

    if (rc == 0) return
    p = afg_last_error(ctx)
    msg = ""
    if (c_associated(p)) then
       n = int(afg_strlen(p))
       call c_f_pointer(p, chars, [n])
       allocate(character(len=n) :: msg)
       do i = 1, n
          msg(i:i) = chars(i)
       end do
    end if
    write(*, "(A,A,A,I0,A,A)") where, ": afivo_amd error ", "", rc, ": ", msg
    error stop "afivo_amd call failed"
  end subroutine afg_check

end module m_afg_capi
