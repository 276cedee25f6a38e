! Drop-in replacement of afivo's multigrid module m_a$D_mg, running the FAS
! multigrid cycles on an MI355X through libafivo_amd (include/afivo_amd.h).
!
! The public names, the mg$D_t options type and the procedure interfaces are
! those of the reference module (src/m_aX_mg.f90:6-125), so user programs
! compile unchanged; this module is compiled against the user's afivo tree
! modules (m_a$D_t, m_a$D_gc) and linked in place of the reference m_a$D_mg.
!
! What runs where:
!  * mg$D_fas_fmg / mg$D_fas_vcycle upload the tree structure (when it
!    changed), the cell variables phi, rhs, tmp (and eps) and the boundary
!    values captured from sides_bc, run the whole cycle on the GPU and copy
!    phi, rhs and tmp back into box%cc (unless mg%device_resident is set).
!  * The operator procedures (mg$D_box_lpl, mg$D_auto_op, ...) identify the
!    operator family by procedure identity for the device cycles.  Called
!    directly, they apply their operator to one box on the host (the
!    library's afg_box_* functions), as the reference's do.  A cylindrical
!    tree (coord = a5_cyl) with the auto operators runs mg2_box_clpl / clpld
!    on the device.  Operator sets other than the lpl set, the auto
!    operators and the lpllsf set (user-written operators) stop in
!    mg$D_fas_fmg / mg$D_fas_vcycle: there is no CPU fallback for a cycle.
!  * A user sides_bc is called on the host once per physical boundary face
!    and solve, on the tree's box, to capture the boundary values it writes
!    (they may depend on the geometry, not on phi).
!
! Dimension template: "$D" is replaced by 2 or 3, then cpp runs (the same
! transform afivo applies to its own m_aX_*.f90 sources).
module m_a$D_mg
  use iso_c_binding
  use m_a$D_t
  use m_afg_capi

  implicit none
  private

  integer, parameter, public :: mg_normal_box = 1 !< Normal box
  integer, parameter, public :: mg_lsf_box = 2    !< Box with an internal boundary
  integer, parameter, public :: mg_ceps_box = 3   !< Box with constant eps /= 1
  integer, parameter, public :: mg_veps_box = 4   !< Box with varying eps (on face)

  !> Multigrid options (components of the reference's mg$D_t, plus
  !> device_resident)
  type, public :: mg$D_t
     integer :: i_phi        = -1
     integer :: i_rhs        = -1
     integer :: i_tmp        = -1
     integer :: i_eps        = -1
     integer :: i_lsf        = -1
     integer :: i_bval       = -1
     integer :: n_cycle_down = -1
     integer :: n_cycle_up   = -1
     integer :: n_cycle_base = -1
     logical :: initialized  = .false.
     !> Keep cell data on the device between calls; mg$D_to_host copies it back
     logical :: device_resident = .false.
     procedure(a$D_subr_bc), pointer, nopass   :: sides_bc => null()
     procedure(a$D_subr_rb), pointer, nopass   :: sides_rb => null()
     procedure(mg$D_box_op), pointer, nopass   :: box_op => null()
     procedure(mg$D_box_gsrb), pointer, nopass :: box_gsrb => null()
     procedure(mg$D_box_corr), pointer, nopass :: box_corr => null()
     procedure(mg$D_box_rstr), pointer, nopass :: box_rstr => null()
  end type mg$D_t

  abstract interface
     subroutine mg$D_box_op(box, i_out, mg)
       import
       type(box$D_t), intent(inout) :: box
       type(mg$D_t), intent(in)     :: mg
       integer, intent(in)          :: i_out
     end subroutine mg$D_box_op

     subroutine mg$D_box_gsrb(box, redblack_cntr, mg)
       import
       type(box$D_t), intent(inout) :: box
       type(mg$D_t), intent(in)     :: mg
       integer, intent(in)          :: redblack_cntr
     end subroutine mg$D_box_gsrb

     subroutine mg$D_box_corr(box_p, box_c, mg)
       import
       type(box$D_t), intent(inout) :: box_c
       type(box$D_t), intent(in)    :: box_p
       type(mg$D_t), intent(in)     :: mg
     end subroutine mg$D_box_corr

     subroutine mg$D_box_rstr(box_c, box_p, iv, mg)
       import
       type(box$D_t), intent(in)    :: box_c
       type(box$D_t), intent(inout) :: box_p
       integer, intent(in)          :: iv
       type(mg$D_t), intent(in)     :: mg
     end subroutine mg$D_box_rstr
  end interface

  public :: mg$D_init_mg
  public :: mg$D_fas_fmg
  public :: mg$D_fas_vcycle
  public :: mg$D_box_op
  public :: mg$D_box_gsrb
  public :: mg$D_box_corr
  public :: mg$D_box_rstr
  public :: mg$D_set_box_tag
  public :: mg$D_auto_op
  public :: mg$D_auto_gsrb
  public :: mg$D_auto_corr
  public :: mg$D_auto_rstr
  public :: mg$D_box_lpl
  public :: mg$D_box_gsrb_lpl
  public :: mg$D_box_corr_lpl
  public :: mg$D_box_rstr_lpl
  public :: mg$D_sides_rb
  public :: mg$D_box_lpld
  public :: mg$D_box_gsrb_lpld
  public :: mg$D_box_corr_lpld
  public :: mg$D_box_lpllsf
  public :: mg$D_box_gsrb_lpllsf
  public :: mg$D_box_corr_lpllsf
  public :: mg$D_box_rstr_lpllsf
#if $D == 2
  public :: mg$D_box_clpl
  public :: mg$D_box_gsrb_clpl
  public :: mg$D_box_clpld
  public :: mg$D_box_gsrb_clpld
#endif
  ! device helpers and extensions (not in the reference)
  public :: mg$D_to_host
  public :: mg$D_device_release
  public :: mg$D_mixed_solve
  public :: mg$D_autotune
  public :: mg$D_device_update
  public :: mg$D_to_device

  ! device context of this module and the structure it holds
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: ctx_nc = -1, ctx_nvar = -1, ctx_coord = -1
  integer(c_int32_t), allocatable, save :: recs_dev(:)
  integer(c_int32_t), allocatable, save :: ids_dev(:)
  integer(c_int64_t), save :: n_values = 0

  integer, parameter :: rec_len = 3 + $D + 2**$D + 2*$D

contains

  !> mg$D_init_mg (m_aX_mg.f90:130-154): check the options and set defaults
  subroutine mg$D_init_mg(mg)
    type(mg$D_t), intent(inout) :: mg

    if (mg%i_phi < 0)                  stop "mg$D_init_mg: i_phi not set"
    if (mg%i_tmp < 0)                  stop "mg$D_init_mg: i_tmp not set"
    if (mg%i_rhs < 0)                  stop "mg$D_init_mg: i_rhs not set"
    if (mg%i_lsf * mg%i_bval < 0) &
         stop "mg$D_init_mg: you have to set both i_lsf and i_bval"
    if (.not. associated(mg%sides_bc)) stop "mg$D_init_mg: sides_bc not set"

    if (mg%n_cycle_down < 0)           mg%n_cycle_down = 2
    if (mg%n_cycle_up < 0)             mg%n_cycle_up = 2
    if (mg%n_cycle_base < 0)           mg%n_cycle_base = 2

    if (.not. associated(mg%sides_rb)) mg%sides_rb => mg$D_sides_rb
    if (.not. associated(mg%box_op))   mg%box_op => mg$D_box_lpl
    if (.not. associated(mg%box_gsrb)) mg%box_gsrb => mg$D_box_gsrb_lpl
    if (.not. associated(mg%box_corr)) mg%box_corr => mg$D_box_corr_lpl
    if (.not. associated(mg%box_rstr)) mg%box_rstr => mg$D_box_rstr_lpl

    mg%initialized = .true.
  end subroutine mg$D_init_mg

  subroutine check_mg(mg)
    type(mg$D_t), intent(in) :: mg
    if (.not. mg%initialized) stop "check_mg: you haven't called mg$D_init_mg"
  end subroutine check_mg

  !> mg$D_fas_fmg (m_aX_mg.f90:164-203) on the device
  subroutine mg$D_fas_fmg(tree, mg, set_residual, have_guess)
    type(a$D_t), intent(inout) :: tree
    type(mg$D_t), intent(in)   :: mg
    logical, intent(in)        :: set_residual
    logical, intent(in)        :: have_guess
    type(afg_mg_opts)          :: o

    call check_mg(mg)
    call device_prepare(tree, mg, o)
    ! init_phi_rhs restricts rhs before any other operator has tagged a box
    call check_rstr_tags(tree, o, .not. have_guess)
    call afg_check(afg_fas_fmg(ctx, o, merge(1, 0, set_residual), &
         merge(1, 0, have_guess)), ctx, "mg$D_fas_fmg")
    call mirror_tags(tree, mg, o)
    if (.not. mg%device_resident) call mg$D_to_host(tree, mg)
  end subroutine mg$D_fas_fmg

  !> mg$D_fas_vcycle (m_aX_mg.f90:208-255) on the device
  subroutine mg$D_fas_vcycle(tree, mg, highest_lvl, set_residual)
    type(a$D_t), intent(inout) :: tree
    type(mg$D_t), intent(in)   :: mg
    integer, intent(in)        :: highest_lvl
    logical, intent(in)        :: set_residual
    type(afg_mg_opts)          :: o

    call check_mg(mg)
    call device_prepare(tree, mg, o)
    call check_rstr_tags(tree, o, .false.)
    call afg_check(afg_fas_vcycle(ctx, o, highest_lvl, merge(1, 0, set_residual)), &
         ctx, "mg$D_fas_vcycle")
    call mirror_tags(tree, mg, o)
    if (.not. mg%device_resident) call mg$D_to_host(tree, mg)
  end subroutine mg$D_fas_vcycle

  !> mg$D_auto_rstr stops on a child whose tag is unset (m_aX_mg.f90:639).
  !> The auto operators tag a box before it is restricted, except in
  !> init_phi_rhs (fmg_init); other operators never tag.  Boxes on levels
  !> above the lowest are checked where the reference could meet an unset tag.
  subroutine check_rstr_tags(tree, o, fmg_init)
    type(a$D_t), intent(in)       :: tree
    type(afg_mg_opts), intent(in) :: o
    logical, intent(in)           :: fmg_init
    integer                       :: lvl, i

    if (o%rstr /= AFG_RSTR_AUTO) return
    if (o%op == AFG_OP_AUTO .and. .not. fmg_init) return
    do lvl = lbound(tree%lvls, 1) + 1, tree%highest_lvl
       do i = 1, size(tree%lvls(lvl)%ids)
          if (tree%boxes(tree%lvls(lvl)%ids(i))%tag == a5_init_tag) &
               stop "mg$D_auto_rstr: box_c tag not set"
       end do
    end do
  end subroutine check_rstr_tags

  !> The auto operators set each box's tag at their first call on it
  !> (m_aX_mg.f90:567,602,655); after a device solve the host tree's tags
  !> are set as the reference's solve leaves them.
  subroutine mirror_tags(tree, mg, o)
    type(a$D_t), intent(inout)    :: tree
    type(mg$D_t), intent(in)      :: mg
    type(afg_mg_opts), intent(in) :: o
    integer                       :: lvl, i, id

    if (o%op /= AFG_OP_AUTO) return
    do lvl = lbound(tree%lvls, 1), tree%highest_lvl
       do i = 1, size(tree%lvls(lvl)%ids)
          id = tree%lvls(lvl)%ids(i)
          if (tree%boxes(id)%tag == a5_init_tag) call mg$D_set_box_tag(tree%boxes(id), mg)
       end do
    end do
  end subroutine mirror_tags

  !> Solve to max|rhs - A phi| <= tol * max|rhs| over the leaves by FP64 defect
  !> correction with FP32 FAS multigrid corrections on the device
  !> (afg_mixed_solve: each correction `cycles` FP32 V-cycles, the first an
  !> FMG); phi starts from zero unless have_guess.  No counterpart in the
  !> reference.  Returns the corrections and the relative residual reached.
  subroutine mg$D_mixed_solve(tree, mg, tol, max_iter, have_guess, n_iter, rel_res, cycles)
    type(a$D_t), intent(inout)    :: tree
    type(mg$D_t), intent(in)      :: mg
    real(dp), intent(in)          :: tol
    integer, intent(in)           :: max_iter
    logical, intent(in)           :: have_guess
    integer, intent(out)          :: n_iter
    real(dp), intent(out)         :: rel_res
    integer, intent(in), optional :: cycles
    type(afg_mg_opts)             :: o
    integer(c_int)                :: it, cyc
    real(c_double)                :: rr

    call check_mg(mg)
    cyc = 3
    if (present(cycles)) cyc = cycles
    call device_prepare(tree, mg, o)
    call afg_check(afg_mixed_solve(ctx, o, merge(1, 0, have_guess), real(tol, c_double), &
         int(max_iter, c_int), cyc, it, rr), ctx, "mg$D_mixed_solve")
    n_iter = it
    rel_res = rr
    if (.not. mg%device_resident) call mg$D_to_host(tree, mg)
  end subroutine mg$D_mixed_solve

  !> Let the device library pick its tile shapes and bottom-kernel limit for
  !> this tree from timed trial runs, for the FP64 and the FP32 stages
  !> (afg_autotune; the tree's data is not changed).  No counterpart in the
  !> reference.
  subroutine mg$D_autotune(tree, mg)
    type(a$D_t), intent(inout) :: tree
    type(mg$D_t), intent(in)   :: mg
    type(afg_mg_opts)          :: o

    call check_mg(mg)
    call device_prepare(tree, mg, o)
    call afg_check(afg_autotune(ctx, o, 0, 8, c_null_ptr), ctx, "mg$D_autotune")
    call afg_check(afg_autotune(ctx, o, 1, 8, c_null_ptr), ctx, "mg$D_autotune")
  end subroutine mg$D_autotune

  !> After a$D_adjust_refinement of the tree (m_aX_core.f90:452-541): move
  !> the device copy of the tree to the new structure without a round trip
  !> through box%cc (afg_update_tree).  Boxes the update kept keep their
  !> device values; the boxes it added get the variables ivs set from their
  !> parents by a$D_prolong<order> (order 0, 1 or 2; m_aX_prolong.f90) and
  !> zeros elsewhere.  For device-resident solves, where the device holds
  !> the newest phi; set the new boxes' rhs on the host and send it with
  !> mg$D_to_device.  No counterpart in the reference.
  subroutine mg$D_device_update(tree, order, ivs)
    type(a$D_t), intent(in) :: tree
    integer, intent(in)     :: order
    integer, intent(in)     :: ivs(:)
    integer(c_int32_t), allocatable :: recs(:)

    if (.not. c_associated(ctx)) return   ! nothing on the device: the next solve uploads
    call build_records(tree, recs)
    call afg_check(afg_update_tree(ctx, size(recs) / rec_len, recs, tree%dr_base, c_null_ptr, &
         int(order, c_int), int(size(ivs), c_int), int(ivs, c_int32_t)), ctx, "mg$D_device_update")
    recs_dev = recs
    ids_dev = recs(1::rec_len)
  end subroutine mg$D_device_update

  !> Copy variable iv of every box from box%cc to the device (for instance
  !> the rhs of boxes added by a refinement).  No counterpart in the
  !> reference.
  subroutine mg$D_to_device(tree, iv)
    type(a$D_t), intent(in) :: tree
    integer, intent(in)     :: iv
    integer(c_int32_t), allocatable :: recs(:)

    if (.not. c_associated(ctx)) return   ! nothing on the device: the next solve uploads
    call build_records(tree, recs)
    if (size(recs) /= size(recs_dev)) error stop "mg$D_to_device: tree changed (call mg$D_device_update first)"
    if (any(recs /= recs_dev)) error stop "mg$D_to_device: tree changed (call mg$D_device_update first)"
    call upload_var(tree, iv)
  end subroutine mg$D_to_device

  !> Copy phi, rhs and tmp from the device into box%cc
  subroutine mg$D_to_host(tree, mg)
    type(a$D_t), intent(inout) :: tree
    type(mg$D_t), intent(in)   :: mg
    if (.not. c_associated(ctx)) return
    call download_var(tree, mg%i_phi)
    call download_var(tree, mg%i_rhs)
    call download_var(tree, mg%i_tmp)
  end subroutine mg$D_to_host

  !> Free the device context of this module
  subroutine mg$D_device_release()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = afg_ctx_destroy(ctx)
    ctx = c_null_ptr
    ctx_nc = -1
    ctx_nvar = -1
    ctx_coord = -1
    if (allocated(recs_dev)) deallocate(recs_dev)
    if (allocated(ids_dev)) deallocate(ids_dev)
  end subroutine mg$D_device_release

  !> Map the options to the device, (re)build the device tree when its
  !> structure changed, and upload the cell data and boundary values.
  subroutine device_prepare(tree, mg, o)
    type(a$D_t), intent(inout)     :: tree
    type(mg$D_t), intent(in)       :: mg
    type(afg_mg_opts), intent(out) :: o
    integer(c_int32_t), allocatable :: recs(:)
    logical :: new_tree

    o%i_phi = mg%i_phi
    o%i_rhs = mg%i_rhs
    o%i_tmp = mg%i_tmp
    o%i_eps = -1
    o%n_cycle_down = mg%n_cycle_down
    o%n_cycle_up = mg%n_cycle_up
    o%n_cycle_base = mg%n_cycle_base
    o%op = operator_family(mg)
    o%rstr = restriction_kind(mg)
    if (o%op == AFG_OP_AUTO) o%i_eps = mg%i_eps
    o%i_lsf = mg%i_lsf
    o%i_bval = mg%i_bval
    call afg_check(afg_mg_init(o), c_null_ptr, "mg$D_init_mg")

    if (.not. associated(mg%sides_rb, mg$D_sides_rb)) &
         error stop "mg$D: only mg$D_sides_rb is implemented on the device for sides_rb"

    if (c_associated(ctx) .and. (ctx_nc /= tree%n_cell .or. ctx_nvar /= tree%n_var_cell)) &
         call mg$D_device_release()
    if (.not. c_associated(ctx)) then
       call afg_check(afg_ctx_create($D, tree%n_cell, tree%n_var_cell, ctx), &
            c_null_ptr, "afg_ctx_create")
       ctx_nc = tree%n_cell
       ctx_nvar = tree%n_var_cell
    end if

    call build_records(tree, recs)
    new_tree = .true.
    if (allocated(recs_dev)) then
       if (size(recs_dev) == size(recs)) new_tree = any(recs_dev /= recs)
    end if
    if (new_tree) then
       call afg_check(afg_set_tree(ctx, size(recs) / rec_len, recs, tree%dr_base), &
            ctx, "afg_set_tree")
       recs_dev = recs
       ids_dev = recs(1::rec_len)
    end if
    if (tree%coord_t /= ctx_coord) then
       call afg_check(afg_set_coord(ctx, merge(AFG_COORD_CYL, AFG_COORD_XYZ, tree%coord_t == a5_cyl), &
            real(tree%r_base(1), c_double)), ctx, "afg_set_coord")
       ctx_coord = tree%coord_t
    end if

    if (new_tree .or. .not. mg%device_resident) then
       call upload_var(tree, mg%i_phi)
       call upload_var(tree, mg%i_rhs)
       call upload_var(tree, mg%i_tmp)
       if (o%op == AFG_OP_AUTO .and. mg%i_eps > 0) call upload_var(tree, mg%i_eps)
       if (mg%i_lsf > 0 .and. (o%op /= AFG_OP_LPL .or. o%rstr /= AFG_RSTR_LPL)) then
          call upload_var(tree, mg%i_lsf)
          call upload_var(tree, mg%i_bval)
       end if
    end if
    call set_boundaries(tree, mg)
  end subroutine device_prepare

  !> Device operator family from the procedure pointers of mg
  integer(c_int) function operator_family(mg) result(op)
    type(mg$D_t), intent(in) :: mg

    if (associated(mg%box_op, mg$D_box_lpl) .and. &
         associated(mg%box_gsrb, mg$D_box_gsrb_lpl) .and. &
         associated(mg%box_corr, mg$D_box_corr_lpl)) then
       op = AFG_OP_LPL
    else if (associated(mg%box_op, mg$D_auto_op) .and. &
         associated(mg%box_gsrb, mg$D_auto_gsrb) .and. &
         associated(mg%box_corr, mg$D_auto_corr)) then
       op = AFG_OP_AUTO
    else if (associated(mg%box_op, mg$D_box_lpllsf) .and. &
         associated(mg%box_gsrb, mg$D_box_gsrb_lpllsf) .and. &
         associated(mg%box_corr, mg$D_box_corr_lpllsf)) then
       op = AFG_OP_LSF
    else
       error stop "mg$D: operator set not implemented on the device " // &
            "(supported: the lpl set, auto_op/auto_gsrb/auto_corr and the lpllsf set)"
    end if
  end function operator_family

  !> Device restriction from mg%box_rstr
  integer(c_int) function restriction_kind(mg) result(rstr)
    type(mg$D_t), intent(in) :: mg

    if (associated(mg%box_rstr, mg$D_box_rstr_lpl)) then
       rstr = AFG_RSTR_LPL
    else if (associated(mg%box_rstr, mg$D_auto_rstr)) then
       rstr = AFG_RSTR_AUTO
    else if (associated(mg%box_rstr, mg$D_box_rstr_lpllsf)) then
       rstr = AFG_RSTR_LSF
    else
       error stop "mg$D: box_rstr not implemented on the device " // &
            "(supported: mg$D_box_rstr_lpl, mg$D_auto_rstr, mg$D_box_rstr_lpllsf)"
    end if
  end function restriction_kind

  !> One record [id, lvl, ix, parent, children, neighbors] per box in use,
  !> in level order (the layout afg_set_tree takes)
  subroutine build_records(tree, recs)
    type(a$D_t), intent(in) :: tree
    integer(c_int32_t), allocatable, intent(out) :: recs(:)
    integer :: lvl, i, id, n, pos

    n = 0
    do lvl = lbound(tree%lvls, 1), tree%highest_lvl
       n = n + size(tree%lvls(lvl)%ids)
    end do
    allocate(recs(n * rec_len))
    pos = 0
    do lvl = lbound(tree%lvls, 1), tree%highest_lvl
       do i = 1, size(tree%lvls(lvl)%ids)
          id = tree%lvls(lvl)%ids(i)
          recs(pos+1) = id
          recs(pos+2) = tree%boxes(id)%lvl
          recs(pos+3:pos+2+$D) = tree%boxes(id)%ix
          recs(pos+3+$D) = tree%boxes(id)%parent
          recs(pos+4+$D:pos+3+$D+2**$D) = tree%boxes(id)%children
          recs(pos+4+$D+2**$D:pos+rec_len) = tree%boxes(id)%neighbors
          pos = pos + rec_len
       end do
    end do
  end subroutine build_records

  !> box%cc(..., iv) of every box to the device: the library copies each box's
  !> slab from where it lies (no packing here)
  subroutine upload_var(tree, iv)
    type(a$D_t), intent(in), target :: tree
    integer, intent(in)             :: iv
    type(c_ptr), allocatable        :: ptrs(:)

    call box_pointers(tree, iv, ptrs)
    call afg_check(afg_upload_boxes(ctx, iv, size(ids_dev), ids_dev, ptrs), ctx, "afg_upload_boxes")
  end subroutine upload_var

  subroutine download_var(tree, iv)
    type(a$D_t), intent(inout), target :: tree
    integer, intent(in)                :: iv
    type(c_ptr), allocatable           :: ptrs(:)

    call box_pointers(tree, iv, ptrs)
    call afg_check(afg_download_boxes(ctx, iv, size(ids_dev), ids_dev, ptrs), ctx, "afg_download_boxes")
  end subroutine download_var

  !> the address of box%cc(0, 0[, 0], iv) of each box of ids_dev: the start of
  !> the box's contiguous array of variable iv
  subroutine box_pointers(tree, iv, ptrs)
    type(a$D_t), intent(in), target       :: tree
    integer, intent(in)                   :: iv
    type(c_ptr), allocatable, intent(out) :: ptrs(:)
    integer                               :: k

    allocate(ptrs(size(ids_dev)))
    do k = 1, size(ids_dev)
#if $D == 2
       ptrs(k) = c_loc(tree%boxes(ids_dev(k))%cc(0, 0, iv))
#elif $D == 3
       ptrs(k) = c_loc(tree%boxes(ids_dev(k))%cc(0, 0, 0, iv))
#endif
    end do
  end subroutine box_pointers


  !> Boundary table: for every physical face, the type and values the user's
  !> sides_bc writes (built-in zero conditions need no call).  sides_bc is
  !> called on the tree's own box, as the reference's ghost fill calls it
  !> (m_aX_gc.f90:58-67): it writes the boundary values into the face's
  !> ghost cells of phi, which are read back here; the ghost fill on the
  !> device turns them into ghost values, and the download of phi after the
  !> solve overwrites them with those.
  subroutine set_boundaries(tree, mg)
    use m_a$D_gc, only: a$D_bc_dirichlet_zero, a$D_bc_neumann_zero
    type(a$D_t), intent(inout) :: tree
    type(mg$D_t), intent(in)   :: mg
    integer                    :: k, id, nb, bc_type
    real(c_double), allocatable, target :: b(:)

    do k = 1, size(ids_dev)
       id = ids_dev(k)
       do nb = 1, a$D_num_neighbors
          if (tree%boxes(id)%neighbors(nb) >= a5_no_box) cycle
          if (associated(mg%sides_bc, a$D_bc_dirichlet_zero)) then
             call afg_check(afg_set_bc(ctx, id, nb, AFG_BC_DIRICHLET, c_null_ptr), ctx, "afg_set_bc")
          else if (associated(mg%sides_bc, a$D_bc_neumann_zero)) then
             call afg_check(afg_set_bc(ctx, id, nb, AFG_BC_NEUMANN, c_null_ptr), ctx, "afg_set_bc")
          else
             call mg%sides_bc(tree%boxes(id), nb, mg%i_phi, bc_type)
             call face_values(tree%boxes(id), nb, mg%i_phi, b)
             call afg_check(afg_set_bc(ctx, id, nb, bc_type, c_loc(b)), ctx, "afg_set_bc")
          end if
       end do
    end do
  end subroutine set_boundaries

  !> Ghost values of face nb, first tangential index fastest
  subroutine face_values(box, nb, iv, b)
    type(box$D_t), intent(in) :: box
    integer, intent(in) :: nb, iv
    real(c_double), allocatable, intent(out) :: b(:)
    integer :: nc, g
    nc = box%n_cell
    g = a$D_neighb_high_01(nb) * (nc + 1)
#if $D == 2
    select case (a$D_neighb_dim(nb))
    case (1)
       b = box%cc(g, 1:nc, iv)
    case (2)
       b = box%cc(1:nc, g, iv)
    end select
#elif $D == 3
    select case (a$D_neighb_dim(nb))
    case (1)
       b = reshape(box%cc(g, 1:nc, 1:nc, iv), [nc*nc])
    case (2)
       b = reshape(box%cc(1:nc, g, 1:nc, iv), [nc*nc])
    case (3)
       b = reshape(box%cc(1:nc, 1:nc, g, iv), [nc*nc])
    end select
#endif
  end subroutine face_values

  !> mg$D_set_box_tag (m_aX_mg.f90:667-706): tag from min/max of eps over the
  !> whole array (ghost cells included) and the sign change of the level set
  subroutine mg$D_set_box_tag(box, mg)
    type(box$D_t), intent(inout) :: box
    type(mg$D_t), intent(in)     :: mg
    real(dp) :: a, b
    logical :: is_lsf, is_deps, is_eps

    is_lsf = .false.
    is_eps = .false.
    is_deps = .false.
#if $D == 2
    if (mg%i_lsf /= -1) is_lsf = minval(box%cc(:, :, mg%i_lsf)) * &
         maxval(box%cc(:, :, mg%i_lsf)) < 0
    if (mg%i_eps /= -1) then
       a = minval(box%cc(:, :, mg%i_eps))
       b = maxval(box%cc(:, :, mg%i_eps))
#elif $D == 3
    if (mg%i_lsf /= -1) is_lsf = minval(box%cc(:, :, :, mg%i_lsf)) * &
         maxval(box%cc(:, :, :, mg%i_lsf)) < 0
    if (mg%i_eps /= -1) then
       a = minval(box%cc(:, :, :, mg%i_eps))
       b = maxval(box%cc(:, :, :, mg%i_eps))
#endif
       is_deps = (b > a)
       if (.not. is_deps) is_eps = (a < 1 .or. a > 1)
    end if
    if (count([is_lsf, is_eps, is_deps]) > 1) &
         stop "mg$D_set_box_tag: Cannot set lsf and eps tag for same box"
    box%tag = mg_normal_box
    if (is_lsf) box%tag = mg_lsf_box
    if (is_eps) box%tag = mg_ceps_box
    if (is_deps) box%tag = mg_veps_box
  end subroutine mg$D_set_box_tag

  ! ------------------------------------------------------------------
  ! Operator procedures.  mg$D_fas_fmg / mg$D_fas_vcycle identify the
  ! operator family by procedure identity and run it on the device.  Called
  ! directly, each applies its operator to the one box on the host, through
  ! the library's per-box operators (afg_box_*: the per-cell code of the
  ! device stages).  Arrays an operator does not read are passed as stand-ins.

#if $D == 2
#define CC(b, iv) b%cc(:, :, iv)
#elif $D == 3
#define CC(b, iv) b%cc(:, :, :, iv)
#endif

  !> iv if it is set, else the stand-in variable
  integer function var_or(iv, stand_in)
    integer, intent(in) :: iv, stand_in
    var_or = merge(iv, stand_in, iv > 0)
  end function var_or

  !> stop unless the variables of operator kind `kind` are set
  subroutine need_vars(mg, kind, name)
    type(mg$D_t), intent(in)     :: mg
    integer, intent(in)          :: kind
    character(len=*), intent(in) :: name
    logical                      :: ok

    ok = .true.
    if (kind == AFG_BOX_LPLD .or. kind == AFG_BOX_CLPLD) ok = mg%i_eps > 0
    if (kind == AFG_BOX_LSF) ok = mg%i_lsf > 0 .and. mg%i_bval > 0
    if (.not. ok) then
       write(*, "(A,A)") name, ": needs mg%i_eps (lpld operators) or mg%i_lsf and mg%i_bval (lpllsf operators)"
       error stop "afivo_amd: multigrid variable not set"
    end if
  end subroutine need_vars

  !> whether the box's operators are the cylindrical ones (2D, a5_cyl)
  logical function is_cyl(box)
    type(box$D_t), intent(in) :: box
#if $D == 2
    is_cyl = box%coord_t == a5_cyl
#else
    is_cyl = .false.
#endif
  end function is_cyl

  subroutine host_op(box, i_out, mg, kind, name)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: i_out, kind
    type(mg$D_t), intent(in)     :: mg
    character(len=*), intent(in) :: name
    integer                      :: ie, il, ib

    call need_vars(mg, kind, name)
    ie = var_or(mg%i_eps, mg%i_phi)
    il = var_or(mg%i_lsf, mg%i_phi)
    ib = var_or(mg%i_bval, mg%i_phi)
    call afg_check(afg_box_op($D, box%n_cell, box%dr, kind, (box%ix(1) - 1) * box%n_cell, &
         CC(box, mg%i_phi), CC(box, ie), CC(box, il), CC(box, ib), CC(box, i_out)), c_null_ptr, name)
  end subroutine host_op

  subroutine host_gsrb(box, redblack_cntr, mg, kind, name)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: redblack_cntr, kind
    type(mg$D_t), intent(in)     :: mg
    character(len=*), intent(in) :: name
    integer                      :: ie, il, ib

    call need_vars(mg, kind, name)
    ie = var_or(mg%i_eps, mg%i_rhs)
    il = var_or(mg%i_lsf, mg%i_rhs)
    ib = var_or(mg%i_bval, mg%i_rhs)
    call afg_check(afg_box_gsrb($D, box%n_cell, box%dr, kind, (box%ix(1) - 1) * box%n_cell, redblack_cntr, &
         CC(box, mg%i_phi), CC(box, mg%i_rhs), CC(box, ie), CC(box, il), CC(box, ib)), c_null_ptr, name)
  end subroutine host_gsrb

  !> the child's phi corrected with the parent's tmp (phi - tmp after
  !> correct_children's subtraction)
  subroutine host_corr(box_p, box_c, mg, kind, name)
    type(box$D_t), intent(inout) :: box_c
    type(box$D_t), intent(in)    :: box_p
    type(mg$D_t), intent(in)     :: mg
    integer, intent(in)          :: kind
    character(len=*), intent(in) :: name
    integer(c_int32_t)           :: off($D)
    integer                      :: ie, il, ic

    call need_vars(mg, kind, name)
    off = a$D_get_child_offset(box_c)
    ie = var_or(mg%i_eps, mg%i_tmp)
    il = var_or(mg%i_lsf, mg%i_tmp)
    ic = var_or(mg%i_lsf, mg%i_rhs)
    call afg_check(afg_box_corr($D, box_c%n_cell, box_p%n_cell, kind, off, CC(box_p, mg%i_tmp), &
         CC(box_p, ie), CC(box_p, il), CC(box_c, mg%i_phi), CC(box_c, ic)), c_null_ptr, name)
  end subroutine host_corr

  subroutine host_rstr(box_c, box_p, iv, mg, lsf, name)
    type(box$D_t), intent(in)    :: box_c
    type(box$D_t), intent(inout) :: box_p
    integer, intent(in)          :: iv
    type(mg$D_t), intent(in)     :: mg
    logical, intent(in)          :: lsf
    character(len=*), intent(in) :: name
    integer(c_int32_t)           :: off($D)
    real(c_double)               :: p_rmin

    off = a$D_get_child_offset(box_c)
    p_rmin = box_p%r_min(1)
    if (lsf) then
       call need_vars(mg, AFG_BOX_LSF, name)
       call afg_check(afg_box_rstr($D, box_c%n_cell, box_p%n_cell, AFG_BOX_LSF, off, CC(box_c, iv), &
            CC(box_c, mg%i_lsf), CC(box_p, mg%i_lsf), p_rmin, box_c%dr, CC(box_p, iv)), c_null_ptr, name)
    else
       ! a$D_restrict_box: the radial weights on a cylindrical parent
       call afg_check(afg_box_rstr($D, box_c%n_cell, box_p%n_cell, &
            merge(AFG_BOX_CLPL, AFG_BOX_LPL, is_cyl(box_p)), off, CC(box_c, iv), CC(box_c, iv), &
            CC(box_c, iv), p_rmin, box_c%dr, CC(box_p, iv)), c_null_ptr, name)
    end if
  end subroutine host_rstr

  !> operator family of a tagged box under the auto operators
  !> (m_aX_mg.f90:562-627); -1 for a tag the auto operators do not handle
  integer function auto_kind(box)
    type(box$D_t), intent(in) :: box

    select case (box%tag)
    case (mg_normal_box)
       auto_kind = merge(AFG_BOX_CLPL, AFG_BOX_LPL, is_cyl(box))
    case (mg_lsf_box)
       auto_kind = AFG_BOX_LSF
    case (mg_veps_box, mg_ceps_box)
       auto_kind = merge(AFG_BOX_CLPLD, AFG_BOX_LPLD, is_cyl(box))
    case default
       auto_kind = -1
    end select
  end function auto_kind

  !> mg$D_sides_rb (m_aX_mg.f90:262-350): ghost cells of face nb at a
  !> refinement boundary, from the parent (a$D_gc_prolong0) and the box's
  !> own cells
  subroutine mg$D_sides_rb(boxes, id, nb, iv)
    type(box$D_t), intent(inout) :: boxes(:)
    integer, intent(in) :: id, nb, iv
    integer(c_int32_t)  :: off($D)
    integer             :: p

    p = boxes(id)%parent
    if (p <= a5_no_box) error stop "mg$D_sides_rb: box has no parent"
    off = a$D_get_child_offset(boxes(id))
    call afg_check(afg_box_sides_rb($D, boxes(id)%n_cell, boxes(p)%n_cell, nb, off, CC(boxes(p), iv), &
         CC(boxes(id), iv)), c_null_ptr, "mg$D_sides_rb")
  end subroutine mg$D_sides_rb

  !> mg$D_box_lpl (m_aX_mg.f90:763-797)
  subroutine mg$D_box_lpl(box, i_out, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: i_out
    type(mg$D_t), intent(in)     :: mg
    call host_op(box, i_out, mg, AFG_BOX_LPL, "mg$D_box_lpl")
  end subroutine mg$D_box_lpl

  !> mg$D_box_gsrb_lpl (m_aX_mg.f90:718-760)
  subroutine mg$D_box_gsrb_lpl(box, redblack_cntr, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: redblack_cntr
    type(mg$D_t), intent(in)     :: mg
    call host_gsrb(box, redblack_cntr, mg, AFG_BOX_LPL, "mg$D_box_gsrb_lpl")
  end subroutine mg$D_box_gsrb_lpl

  !> mg$D_box_corr_lpl (m_aX_mg.f90:708-715)
  subroutine mg$D_box_corr_lpl(box_p, box_c, mg)
    type(box$D_t), intent(inout) :: box_c
    type(box$D_t), intent(in)    :: box_p
    type(mg$D_t), intent(in)     :: mg
    call host_corr(box_p, box_c, mg, AFG_BOX_LPL, "mg$D_box_corr_lpl")
  end subroutine mg$D_box_corr_lpl

  !> mg$D_box_rstr_lpl (m_aX_mg.f90:800-808)
  subroutine mg$D_box_rstr_lpl(box_c, box_p, iv, mg)
    type(box$D_t), intent(in)    :: box_c
    type(box$D_t), intent(inout) :: box_p
    integer, intent(in)          :: iv
    type(mg$D_t), intent(in)     :: mg
    call host_rstr(box_c, box_p, iv, mg, .false., "mg$D_box_rstr_lpl")
  end subroutine mg$D_box_rstr_lpl

  !> mg$D_auto_op (m_aX_mg.f90:597-629)
  subroutine mg$D_auto_op(box, i_out, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: i_out
    type(mg$D_t), intent(in)     :: mg
    if (box%tag == a5_init_tag) call mg$D_set_box_tag(box, mg)
    if (auto_kind(box) >= 0) call host_op(box, i_out, mg, auto_kind(box), "mg$D_auto_op")
  end subroutine mg$D_auto_op

  !> mg$D_auto_gsrb (m_aX_mg.f90:562-594)
  subroutine mg$D_auto_gsrb(box, redblack_cntr, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: redblack_cntr
    type(mg$D_t), intent(in)     :: mg
    if (box%tag == a5_init_tag) call mg$D_set_box_tag(box, mg)
    if (auto_kind(box) >= 0) call host_gsrb(box, redblack_cntr, mg, auto_kind(box), "mg$D_auto_gsrb")
  end subroutine mg$D_auto_gsrb

  !> mg$D_auto_corr (m_aX_mg.f90:650-665): by the child's tag, corr_lpl on
  !> cylindrical trees too
  subroutine mg$D_auto_corr(box_p, box_c, mg)
    type(box$D_t), intent(inout) :: box_c
    type(box$D_t), intent(in)    :: box_p
    type(mg$D_t), intent(in)     :: mg
    if (box_c%tag == a5_init_tag) call mg$D_set_box_tag(box_c, mg)
    select case (box_c%tag)
    case (mg_normal_box)
       call host_corr(box_p, box_c, mg, AFG_BOX_LPL, "mg$D_auto_corr")
    case (mg_lsf_box)
       call host_corr(box_p, box_c, mg, AFG_BOX_LSF, "mg$D_auto_corr")
    case (mg_veps_box, mg_ceps_box)
       call host_corr(box_p, box_c, mg, AFG_BOX_LPLD, "mg$D_auto_corr")
    end select
  end subroutine mg$D_auto_corr

  !> mg$D_auto_rstr (m_aX_mg.f90:632-647)
  subroutine mg$D_auto_rstr(box_c, box_p, iv, mg)
    type(box$D_t), intent(in)    :: box_c
    type(box$D_t), intent(inout) :: box_p
    integer, intent(in)          :: iv
    type(mg$D_t), intent(in)     :: mg
    if (box_c%tag == a5_init_tag) stop "mg$D_auto_rstr: box_c tag not set"
    select case (box_c%tag)
    case (mg_normal_box, mg_veps_box, mg_ceps_box)
       call host_rstr(box_c, box_p, iv, mg, .false., "mg$D_auto_rstr")
    case (mg_lsf_box)
       call host_rstr(box_c, box_p, iv, mg, .true., "mg$D_auto_rstr")
    end select
  end subroutine mg$D_auto_rstr

  !> mg$D_box_lpld (m_aX_mg.f90:866-916)
  subroutine mg$D_box_lpld(box, i_out, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: i_out
    type(mg$D_t), intent(in)     :: mg
    call host_op(box, i_out, mg, AFG_BOX_LPLD, "mg$D_box_lpld")
  end subroutine mg$D_box_lpld

  !> mg$D_box_gsrb_lpld (m_aX_mg.f90:810-863)
  subroutine mg$D_box_gsrb_lpld(box, redblack_cntr, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: redblack_cntr
    type(mg$D_t), intent(in)     :: mg
    call host_gsrb(box, redblack_cntr, mg, AFG_BOX_LPLD, "mg$D_box_gsrb_lpld")
  end subroutine mg$D_box_gsrb_lpld

  !> mg$D_box_corr_lpld (m_aX_mg.f90:918-985)
  subroutine mg$D_box_corr_lpld(box_p, box_c, mg)
    type(box$D_t), intent(inout) :: box_c
    type(box$D_t), intent(in)    :: box_p
    type(mg$D_t), intent(in)     :: mg
    call host_corr(box_p, box_c, mg, AFG_BOX_LPLD, "mg$D_box_corr_lpld")
  end subroutine mg$D_box_corr_lpld

  !> mg$D_box_lpllsf (m_aX_mg.f90:1183-1253)
  subroutine mg$D_box_lpllsf(box, i_out, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: i_out
    type(mg$D_t), intent(in)     :: mg
    call host_op(box, i_out, mg, AFG_BOX_LSF, "mg$D_box_lpllsf")
  end subroutine mg$D_box_lpllsf

  !> mg$D_box_gsrb_lpllsf (m_aX_mg.f90:1103-1180)
  subroutine mg$D_box_gsrb_lpllsf(box, redblack_cntr, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: redblack_cntr
    type(mg$D_t), intent(in)     :: mg
    call host_gsrb(box, redblack_cntr, mg, AFG_BOX_LSF, "mg$D_box_gsrb_lpllsf")
  end subroutine mg$D_box_gsrb_lpllsf

  !> mg$D_box_corr_lpllsf (m_aX_mg.f90:1020-1100)
  subroutine mg$D_box_corr_lpllsf(box_p, box_c, mg)
    type(box$D_t), intent(inout) :: box_c
    type(box$D_t), intent(in)    :: box_p
    type(mg$D_t), intent(in)     :: mg
    call host_corr(box_p, box_c, mg, AFG_BOX_LSF, "mg$D_box_corr_lpllsf")
  end subroutine mg$D_box_corr_lpllsf

  !> mg$D_box_rstr_lpllsf (m_aX_mg.f90:1256-1321)
  subroutine mg$D_box_rstr_lpllsf(box_c, box_p, iv, mg)
    type(box$D_t), intent(in)    :: box_c
    type(box$D_t), intent(inout) :: box_p
    integer, intent(in)          :: iv
    type(mg$D_t), intent(in)     :: mg
    call host_rstr(box_c, box_p, iv, mg, .true., "mg$D_box_rstr_lpllsf")
  end subroutine mg$D_box_rstr_lpllsf

#if $D == 2
  !> mg2_box_clpl (m_aX_mg.f90:1354-1376)
  subroutine mg$D_box_clpl(box, i_out, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: i_out
    type(mg$D_t), intent(in)     :: mg
    call host_op(box, i_out, mg, AFG_BOX_CLPL, "mg$D_box_clpl")
  end subroutine mg$D_box_clpl

  !> mg2_box_gsrb_clpl (m_aX_mg.f90:1325-1351)
  subroutine mg$D_box_gsrb_clpl(box, redblack_cntr, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: redblack_cntr
    type(mg$D_t), intent(in)     :: mg
    call host_gsrb(box, redblack_cntr, mg, AFG_BOX_CLPL, "mg$D_box_gsrb_clpl")
  end subroutine mg$D_box_gsrb_clpl

  !> mg2_box_clpld (m_aX_mg.f90:1379-1407)
  subroutine mg$D_box_clpld(box, i_out, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: i_out
    type(mg$D_t), intent(in)     :: mg
    call host_op(box, i_out, mg, AFG_BOX_CLPLD, "mg$D_box_clpld")
  end subroutine mg$D_box_clpld

  !> mg2_box_gsrb_clpld (m_aX_mg.f90:1411-1443)
  subroutine mg$D_box_gsrb_clpld(box, redblack_cntr, mg)
    type(box$D_t), intent(inout) :: box
    integer, intent(in)          :: redblack_cntr
    type(mg$D_t), intent(in)     :: mg
    call host_gsrb(box, redblack_cntr, mg, AFG_BOX_CLPLD, "mg$D_box_gsrb_clpld")
  end subroutine mg$D_box_gsrb_clpld
#endif

end module m_a$D_mg
