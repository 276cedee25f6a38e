! Test harness that drives the reference afivo multigrid solver (built from the
! reference sources into oracle/_ref by oracle/Makefile) on a configured tree
! and writes the tree and the cell data to a binary dump.  This file is test
! infrastructure: it is the checker, not the product.
!
! Dimension template: "$D" is replaced by 2 or 3 and the result is passed
! through cpp, the same transform the reference applies to its own m_aX_*.f90
! files (src/Makefile:28-32).
!
! Problem set-up follows the reference examples:
!   tree build loop      examples/poisson_basic_2d.f90:41-68
!   Gaussian rhs         examples/m_gaussians.f90:90-102 (gauss_lpl)
!   zero initialisation  all cc set to 0 before the rhs (the reference does not
!                        initialise cc, m_aX_core.f90:361-369)
!   solve                mg$D_fas_fmg / mg$D_fas_vcycle (m_aX_mg.f90:164,208)
!   residual norm        a$D_tree_min_cc / a$D_tree_max_cc (m_aX_utils.f90:468-571)
program ref_harness_$Dd
  use omp_lib
  use m_a$D_t
  use m_a$D_core
  use m_a$D_mg
  use m_a$D_utils
  use m_a$D_gc
  use m_gaussians

  implicit none

  integer, parameter :: i_phi = 1, i_rhs = 2, i_tmp = 3, i_eps = 4
  integer, parameter :: i_lsf = 5, i_bval = 6
  integer            :: n_var = 4
  integer, parameter :: max_gauss = 16

  ! Configuration (namelist /cfg/)
  integer            :: nc = 8
  character(len=16)  :: ref_mode = 'gauss'  ! 'gauss' or 'list'
  integer            :: l_uniform = 3
  integer            :: l_max = 5
  real(dp)           :: thr = 1.0e-4_dp
  integer            :: n_gauss = 1
  real(dp)           :: ampl(max_gauss) = 1.0_dp
  real(dp)           :: sigma(max_gauss) = 0.1_dp
  real(dp)           :: r0($D, max_gauss) = 0.5_dp
  character(len=512) :: list_file = ''
  integer            :: n_down = 2, n_up = 2, n_base = 2
  integer            :: n_fmg = 1
  integer            :: n_vcycle = 0
  logical            :: use_eps = .false.
  real(dp)           :: eps_hi = 1.0_dp
  real(dp)           :: eps_cut = 0.5_dp
  character(len=512) :: out_file = 'ref_out'
  logical            :: dump_input = .true.
  logical            :: dump_fmg = .false.
  logical            :: dump_output = .true.
  integer            :: n_time = 0
  ! refinement step after the solve (0: none): boxes refined below
  ! refine2_lmax where dr^2 max|rhs| > refine2_thr, flagged for removal where
  ! it is below refine2_rm_thr; then the variables 1..refine2_nvar of each
  ! new box set from its parent with a$D_prolong1 (refine2 = 1),
  ! a$D_prolong2 (2) or a$D_prolong0 (3), level by level (as
  ! examples/random_refinement_2d.f90:121-144 does), and the tree written to
  ! out_file.ref2
  integer            :: refine2 = 0
  integer            :: refine2_lmax = 0
  real(dp)           :: refine2_thr = 1.0_dp
  real(dp)           :: refine2_rm_thr = -1.0_dp
  integer            :: refine2_nvar = 1
  ! transfer dumps after the solve (before any refine2 step): variable 4 set
  ! from phi or rhs and then changed by one ghost-cell or transfer routine of
  ! afivo, the tree written after each (see transfer_dumps)
  logical            :: xfer = .false.
  ! cylindrical coordinates (2D: x is the radius, a5_cyl), solved with the
  ! auto operators as examples/poisson_cyl.f90:41-69 does
  logical            :: cyl = .false.
  ! internal boundary given by a level set (variables 5 and 6 added): the
  ! sphere |r - lsf_c| = lsf_rad, with boundary value 1 + x there.
  ! lsf_mode 1: the auto operators (mg_lsf_box tags), with box_rstr =>
  ! mg$D_auto_rstr when rstr_auto (every tag set by mg$D_set_box_tag first);
  ! lsf_mode 2: the lpllsf operator set for every box
  integer            :: lsf_mode = 0
  logical            :: rstr_auto = .false.
  logical            :: tag_boxes = .true.   ! with rstr_auto: mg$D_set_box_tag on every box first
  real(dp)           :: lsf_c($D) = 0.5_dp
  real(dp)           :: lsf_rad = 0.25_dp
  ! builds against the device drop-in only: keep the cell data on the device
  ! between calls (mg%device_resident), copying it back after each solve
  logical            :: resident = .false.

#ifdef AFG_DROPIN
  namelist /cfg/ nc, ref_mode, l_uniform, l_max, thr, n_gauss, ampl, sigma, &
       r0, list_file, n_down, n_up, n_base, n_fmg, n_vcycle, use_eps, eps_hi, &
       eps_cut, out_file, dump_input, dump_fmg, dump_output, n_time, refine2, &
       refine2_lmax, refine2_thr, refine2_rm_thr, refine2_nvar, xfer, cyl, &
       lsf_mode, rstr_auto, tag_boxes, lsf_c, lsf_rad, resident
#else
  namelist /cfg/ nc, ref_mode, l_uniform, l_max, thr, n_gauss, ampl, sigma, &
       r0, list_file, n_down, n_up, n_base, n_fmg, n_vcycle, use_eps, eps_hi, &
       eps_cut, out_file, dump_input, dump_fmg, dump_output, n_time, refine2, &
       refine2_lmax, refine2_thr, refine2_rm_thr, refine2_nvar, xfer, cyl, &
       lsf_mode, rstr_auto, tag_boxes, lsf_c, lsf_rad
#endif

  type(a$D_t)         :: tree, tree_init
  type(ref_info_t)    :: ref_info
  type(mg$D_t)        :: mg
  type(gauss_t)       :: gs
  integer             :: ix_list($D, 1), nb_list(2*$D, 1)
  integer             :: i, n_list, u
  integer, allocatable :: list_key(:, :)
  real(dp)            :: res_hist(0:200), t0, t1
  real(dp), allocatable :: t_runs(:)
  character(len=512)  :: cfg_file

  call get_command_argument(1, cfg_file)
  open(newunit=u, file=trim(cfg_file), status='old', action='read')
  read(u, nml=cfg)
  close(u)

  call gauss_init(gs, ampl(1:n_gauss), sigma(1:n_gauss), r0(:, 1:n_gauss))

  if (ref_mode == 'list') call read_list()
  if (lsf_mode > 0) n_var = 6

#if $D == 2
  call a$D_init(tree, nc, n_var, merge(1, 0, xfer), 1.0_dp / nc, coarsen_to=2, &
       coord=merge(a5_cyl, a5_xyz, cyl), &
#else
  call a$D_init(tree, nc, n_var, merge(1, 0, xfer), 1.0_dp / nc, coarsen_to=2, &
#endif
       mem_limit_gb=64.0_dp)
  ix_list(:, 1) = 1
  nb_list(:, 1) = -1
  call a$D_set_base(tree, ix_list, nb_list)

  do
     if (ref_mode == 'list') then
        call a$D_adjust_refinement(tree, ref_list, ref_info)
     else
        call a$D_adjust_refinement(tree, ref_gauss, ref_info)
     end if
     if (ref_info%n_add == 0) exit
  end do

  call a$D_loop_box(tree, set_init)

  mg%i_phi    = i_phi
  mg%i_rhs    = i_rhs
  mg%i_tmp    = i_tmp
  mg%sides_bc => a$D_bc_dirichlet_zero
  mg%n_cycle_down = n_down
  mg%n_cycle_up   = n_up
  mg%n_cycle_base = n_base
  if (lsf_mode > 0) then
     mg%i_lsf  = i_lsf
     mg%i_bval = i_bval
  end if
  if (lsf_mode == 2) then
     mg%box_op   => mg$D_box_lpllsf
     mg%box_gsrb => mg$D_box_gsrb_lpllsf
     mg%box_corr => mg$D_box_corr_lpllsf
     mg%box_rstr => mg$D_box_rstr_lpllsf
  else if (use_eps .or. cyl .or. lsf_mode == 1) then
     if (use_eps) mg%i_eps = i_eps
     mg%box_op   => mg$D_auto_op
     mg%box_gsrb => mg$D_auto_gsrb
     mg%box_corr => mg$D_auto_corr
     ! box_rstr stays mg$D_box_rstr_lpl, as in
     ! examples/poisson_cyl_dielectric.f90:62-64 (auto_rstr would stop in
     ! init_phi_rhs, m_aX_mg.f90:639, because no tag is set yet)
  end if
  if (rstr_auto) mg%box_rstr => mg$D_auto_rstr
  call mg$D_init_mg(mg)
#ifdef AFG_DROPIN
  mg%device_resident = resident
#endif
  if (rstr_auto .and. tag_boxes) call a$D_loop_box(tree, set_tag)

  if (dump_input) call write_dump(trim(out_file) // '.in')

  if (n_time > 0) tree_init = tree

  call solve(.true.)
#ifdef AFG_DROPIN
  if (resident) call mg$D_to_host(tree, mg)
#endif

  if (dump_output) call write_dump(trim(out_file) // '.out')

  if (xfer) call transfer_dumps()

  if (refine2 > 0) then
     call a$D_adjust_refinement(tree, ref_second, ref_info)
     call prolong_new_boxes()
     call write_dump(trim(out_file) // '.ref2')
  end if

  open(newunit=u, file=trim(out_file) // '.res', status='replace', &
       action='write')
  do i = 0, n_fmg + n_vcycle - 1
     write(u, '(ES25.17E3)') res_hist(i)
  end do
  close(u)

  if (n_time > 0) then
     allocate(t_runs(n_time))
     do i = 1, n_time
        tree = tree_init
        t0 = omp_get_wtime()
        call solve(.false.)
#ifdef AFG_DROPIN
        if (resident) call mg$D_to_host(tree, mg)
#endif
        t1 = omp_get_wtime()
        t_runs(i) = t1 - t0
     end do
     open(newunit=u, file=trim(out_file) // '.time', status='replace', &
          action='write')
     write(u, '(I0)') omp_get_max_threads()
     do i = 1, n_time
        write(u, '(ES25.17E3)') t_runs(i)
     end do
     close(u)
  end if

contains

  !> mg$D_set_box_tag of a box (what mg$D_auto_rstr requires)
  subroutine set_tag(box)
    type(box$D_t), intent(inout) :: box
    call mg$D_set_box_tag(box, mg)
  end subroutine set_tag

  !> FMG (n_fmg times, warm start after the first) then n_vcycle V-cycles,
  !> as in examples/poisson_basic_2d.f90:84-98.
  subroutine solve(record)
    logical, intent(in) :: record
    integer :: it
    real(dp) :: rmin, rmax

    do it = 1, n_fmg
       call mg$D_fas_fmg(tree, mg, .true., it > 1)
       if (record) then
          call a$D_tree_min_cc(tree, i_tmp, rmin)
          call a$D_tree_max_cc(tree, i_tmp, rmax)
          res_hist(it-1) = max(abs(rmin), abs(rmax))
          if (dump_fmg .and. it == 1) call write_dump(trim(out_file) // '.fmg')
       end if
    end do
    do it = 1, n_vcycle
       call mg$D_fas_vcycle(tree, mg, tree%highest_lvl, .true.)
       if (record) then
          call a$D_tree_min_cc(tree, i_tmp, rmin)
          call a$D_tree_max_cc(tree, i_tmp, rmax)
          res_hist(n_fmg+it-1) = max(abs(rmin), abs(rmax))
       end if
    end do
  end subroutine solve

  !> Refine below l_uniform everywhere, and below l_max where
  !> dr^2 * max|rhs| > thr over the box interior.
  subroutine ref_gauss(boxes, id, ref_flag)
    type(box$D_t), intent(in) :: boxes(:)
    integer, intent(in)       :: id
    integer, intent(inout)    :: ref_flag
    integer                   :: i, j, nc_b
#if $D == 3
    integer                   :: k
#endif
    real(dp)                  :: dr2, rmax

    if (boxes(id)%lvl < l_uniform) then
       ref_flag = a5_do_ref
       return
    end if
    if (boxes(id)%lvl >= l_max) return

    nc_b = boxes(id)%n_cell
    dr2 = boxes(id)%dr**2
    rmax = 0
#if $D == 2
    do j = 1, nc_b
       do i = 1, nc_b
          rmax = max(rmax, abs(gauss_lpl(gs, a$D_r_cc(boxes(id), [i, j]))))
       end do
    end do
#elif $D == 3
    do k = 1, nc_b
       do j = 1, nc_b
          do i = 1, nc_b
             rmax = max(rmax, abs(gauss_lpl(gs, &
                  a$D_r_cc(boxes(id), [i, j, k]))))
          end do
       end do
    end do
#endif
    if (dr2 * rmax > thr) ref_flag = a5_do_ref
  end subroutine ref_gauss

  !> The refinement rule of the step after the solve (see refine2_*)
  subroutine ref_second(boxes, id, ref_flag)
    type(box$D_t), intent(in) :: boxes(:)
    integer, intent(in)       :: id
    integer, intent(inout)    :: ref_flag
    integer                   :: i, j, nc_b
#if $D == 3
    integer                   :: k
#endif
    real(dp)                  :: dr2, rmax

    if (boxes(id)%lvl < l_uniform) then
       ref_flag = a5_do_ref
       return
    end if
    nc_b = boxes(id)%n_cell
    dr2 = boxes(id)%dr**2
    rmax = 0
#if $D == 2
    do j = 1, nc_b
       do i = 1, nc_b
          rmax = max(rmax, abs(gauss_lpl(gs, a$D_r_cc(boxes(id), [i, j]))))
       end do
    end do
#elif $D == 3
    do k = 1, nc_b
       do j = 1, nc_b
          do i = 1, nc_b
             rmax = max(rmax, abs(gauss_lpl(gs, &
                  a$D_r_cc(boxes(id), [i, j, k]))))
          end do
       end do
    end do
#endif
    if (boxes(id)%lvl < refine2_lmax .and. dr2 * rmax > refine2_thr) then
       ref_flag = a5_do_ref
    else if (dr2 * rmax < refine2_rm_thr) then
       ref_flag = a5_rm_ref
    end if
  end subroutine ref_second

  !> Variables 1..refine2_nvar of every box added by the step after the
  !> solve, from its parent, lowest level first
  subroutine prolong_new_boxes()
    use m_a$D_prolong
    integer :: lvl, i, id, iv

    do lvl = lbound(ref_info%lvls, 1), ubound(ref_info%lvls, 1)
       if (.not. allocated(ref_info%lvls(lvl)%add)) cycle
       do i = 1, size(ref_info%lvls(lvl)%add)
          id = ref_info%lvls(lvl)%add(i)
          do iv = 1, refine2_nvar
             if (refine2 == 1) then
                call a$D_prolong1(tree%boxes(tree%boxes(id)%parent), tree%boxes(id), iv)
             else if (refine2 == 2) then
                call a$D_prolong2(tree%boxes(tree%boxes(id)%parent), tree%boxes(id), iv)
             else
                call a$D_prolong0(tree%boxes(tree%boxes(id)%parent), tree%boxes(id), iv)
             end if
          end do
       end do
    end do
  end subroutine prolong_new_boxes

  !> Ghost cells and transfers of afivo on variable 4 (a$D_tree_copy_cc from
  !> phi or rhs first), one dump each:
  !>   .gci  a$D_gc_tree with a$D_gc_interp and a$D_bc_dirichlet_zero
  !>   .gcl  a$D_gc_tree with a$D_gc_interp_lim and a$D_bc_neumann_zero
  !>   .gc0  a$D_gc_tree with a$D_gc_prolong0 and a$D_bc_neumann_zero
  !>   .rst  a$D_restrict_tree of a copy of rhs
  !>   .pr1  a$D_prolong1_from(add) of phi into a copy of rhs, every parent
  !>   .pr0  a$D_prolong0_from of phi into a copy of rhs, every parent
  !>   .pr2  a$D_prolong2_from of phi into a copy of rhs, every parent
  !>   .flx  face variable 1 (the gradient of phi) after a$D_consistent_fluxes
  subroutine transfer_dumps()
    use m_a$D_restrict
    use m_a$D_prolong
    integer, parameter :: i_x = 4
    integer :: lvl, i

    call a$D_tree_copy_cc(tree, i_phi, i_x)
    call a$D_gc_tree(tree, i_x, a$D_gc_interp, a$D_bc_dirichlet_zero)
    call write_dump(trim(out_file) // '.gci')
    call a$D_tree_copy_cc(tree, i_phi, i_x)
    call a$D_gc_tree(tree, i_x, a$D_gc_interp_lim, a$D_bc_neumann_zero)
    call write_dump(trim(out_file) // '.gcl')
    call a$D_tree_copy_cc(tree, i_phi, i_x)
    call a$D_gc_tree(tree, i_x, a$D_gc_prolong0, a$D_bc_neumann_zero)
    call write_dump(trim(out_file) // '.gc0')
    call write_gc2(trim(out_file) // '.gc2', i_x)
    call a$D_tree_copy_cc(tree, i_rhs, i_x)
    call a$D_restrict_tree(tree, i_x)
    call write_dump(trim(out_file) // '.rst')
    call a$D_tree_copy_cc(tree, i_rhs, i_x)
    do lvl = lbound(tree%lvls, 1), tree%highest_lvl - 1
       do i = 1, size(tree%lvls(lvl)%parents)
          call a$D_prolong1_from(tree%boxes, tree%lvls(lvl)%parents(i), i_phi, i_x, .true.)
       end do
    end do
    call write_dump(trim(out_file) // '.pr1')
    call a$D_tree_copy_cc(tree, i_rhs, i_x)
    do lvl = lbound(tree%lvls, 1), tree%highest_lvl - 1
       do i = 1, size(tree%lvls(lvl)%parents)
          call a$D_prolong0_from(tree%boxes, tree%lvls(lvl)%parents(i), i_phi, i_x)
       end do
    end do
    call write_dump(trim(out_file) // '.pr0')
    call a$D_tree_copy_cc(tree, i_rhs, i_x)
    do lvl = lbound(tree%lvls, 1), tree%highest_lvl - 1
       do i = 1, size(tree%lvls(lvl)%parents)
          call a$D_prolong2_from(tree%boxes, tree%lvls(lvl)%parents(i), i_phi, i_x)
       end do
    end do
    call write_dump(trim(out_file) // '.pr2')
    ! face variable 1 = the gradient of phi, (phi(i) - phi(i-1)) / dr on
    ! every face, then a$D_consistent_fluxes
    call a$D_loop_box(tree, set_gradient)
    call a$D_consistent_fluxes(tree, [1])
    call write_faces(trim(out_file) // '.flx')
  end subroutine transfer_dumps

  subroutine set_gradient(box)
    type(box$D_t), intent(inout) :: box
    integer :: nc_b
    nc_b = box%n_cell
#if $D == 2
    box%fx(:, :, 1) = (box%cc(1:nc_b+1, 1:nc_b, i_phi) - box%cc(0:nc_b, 1:nc_b, i_phi)) / box%dr
    box%fy(:, :, 1) = (box%cc(1:nc_b, 1:nc_b+1, i_phi) - box%cc(1:nc_b, 0:nc_b, i_phi)) / box%dr
#elif $D == 3
    box%fx(:, :, :, 1) = (box%cc(1:nc_b+1, 1:nc_b, 1:nc_b, i_phi) - &
         box%cc(0:nc_b, 1:nc_b, 1:nc_b, i_phi)) / box%dr
    box%fy(:, :, :, 1) = (box%cc(1:nc_b, 1:nc_b+1, 1:nc_b, i_phi) - &
         box%cc(1:nc_b, 0:nc_b, 1:nc_b, i_phi)) / box%dr
    box%fz(:, :, :, 1) = (box%cc(1:nc_b, 1:nc_b, 1:nc_b+1, i_phi) - &
         box%cc(1:nc_b, 1:nc_b, 0:nc_b, i_phi)) / box%dr
#endif
  end subroutine set_gradient

  !> a$D_gc2_box of variable iv (a$D_gc2_prolong1 at refinement boundaries,
  !> a$D_bc2_neumann_zero at physical ones) for every box, boxes in the order
  !> of write_dump: gc_data(nc[, nc], 2*$D) of each
  subroutine write_gc2(fname, iv)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: iv
    integer :: lu, lvl, n, id, nc_b
#if $D == 2
    real(dp), allocatable :: gc(:, :)
#elif $D == 3
    real(dp), allocatable :: gc(:, :, :)
#endif

    open(newunit=lu, file=fname, access='stream', form='unformatted', &
         status='replace', action='write')
    do lvl = lbound(tree%lvls, 1), tree%highest_lvl
       do n = 1, size(tree%lvls(lvl)%ids)
          id = tree%lvls(lvl)%ids(n)
          nc_b = tree%boxes(id)%n_cell
#if $D == 2
          allocate(gc(nc_b, 2*$D))
#elif $D == 3
          allocate(gc(nc_b, nc_b, 2*$D))
#endif
          call a$D_gc2_box(tree%boxes, id, iv, a$D_gc2_prolong1, a$D_bc2_neumann_zero, gc, nc_b)
          write(lu) gc
          deallocate(gc)
       end do
    end do
    close(lu)
  end subroutine write_gc2

  !> Face variable 1 of every box, boxes in the order of write_dump: fx, fy
  !> (, fz), each in Fortran order
  subroutine write_faces(fname)
    character(len=*), intent(in) :: fname
    integer :: lu, lvl, n, id

    open(newunit=lu, file=fname, access='stream', form='unformatted', &
         status='replace', action='write')
    do lvl = lbound(tree%lvls, 1), tree%highest_lvl
       do n = 1, size(tree%lvls(lvl)%ids)
          id = tree%lvls(lvl)%ids(n)
#if $D == 2
          write(lu) tree%boxes(id)%fx(:, :, 1), tree%boxes(id)%fy(:, :, 1)
#elif $D == 3
          write(lu) tree%boxes(id)%fx(:, :, :, 1), tree%boxes(id)%fy(:, :, :, 1), &
               tree%boxes(id)%fz(:, :, :, 1)
#endif
       end do
    end do
    close(lu)
  end subroutine write_faces

  !> Refine exactly the boxes listed in list_file (sorted by lvl, ix($D) ..
  !> ix(1), lexicographically).
  subroutine ref_list(boxes, id, ref_flag)
    type(box$D_t), intent(in) :: boxes(:)
    integer, intent(in)       :: id
    integer, intent(inout)    :: ref_flag
    integer                   :: key($D+1), lo, hi, mid, c

    key(1) = boxes(id)%lvl
    do c = 1, $D
       key(1+c) = boxes(id)%ix($D+1-c)
    end do
    lo = 1
    hi = n_list
    do while (lo <= hi)
       mid = (lo + hi) / 2
       c = key_cmp(list_key(:, mid), key)
       if (c == 0) then
          ref_flag = a5_do_ref
          return
       else if (c < 0) then
          lo = mid + 1
       else
          hi = mid - 1
       end if
    end do
  end subroutine ref_list

  integer function key_cmp(a, b)
    integer, intent(in) :: a(:), b(:)
    integer :: n
    key_cmp = 0
    do n = 1, size(a)
       if (a(n) < b(n)) then
          key_cmp = -1
          return
       else if (a(n) > b(n)) then
          key_cmp = 1
          return
       end if
    end do
  end function key_cmp

  subroutine read_list()
    integer :: lu, n
    open(newunit=lu, file=trim(list_file), status='old', action='read')
    read(lu, *) n_list
    allocate(list_key($D+1, n_list))
    do n = 1, n_list
       read(lu, *) list_key(:, n)
    end do
    close(lu)
  end subroutine read_list

  !> Zero every variable (including ghost cells), then set rhs = gauss_lpl on
  !> the interior and, if requested, eps on interior and ghost cells: eps = 1
  !> below eps_cut in the last coordinate and eps_hi above.
  subroutine set_init(box)
    type(box$D_t), intent(inout) :: box
    integer                      :: i, j, nc_b
#if $D == 3
    integer                      :: k
#endif
    real(dp)                     :: r($D)

    nc_b = box%n_cell
    box%cc = 0.0_dp
#if $D == 2
    do j = 1, nc_b
       do i = 1, nc_b
          box%cc(i, j, i_rhs) = gauss_lpl(gs, a$D_r_cc(box, [i, j]))
       end do
    end do
    if (lsf_mode > 0) then
       do j = 0, nc_b+1
          do i = 0, nc_b+1
             r = a$D_r_cc(box, [i, j])
             box%cc(i, j, i_lsf) = norm2(r - lsf_c) - lsf_rad
             box%cc(i, j, i_bval) = 1.0_dp + r(1)
          end do
       end do
    end if
    if (use_eps) then
       do j = 0, nc_b+1
          do i = 0, nc_b+1
             r = a$D_r_cc(box, [i, j])
             if (r($D) < eps_cut) then
                box%cc(i, j, i_eps) = 1.0_dp
             else
                box%cc(i, j, i_eps) = eps_hi
             end if
          end do
       end do
    end if
#elif $D == 3
    do k = 1, nc_b
       do j = 1, nc_b
          do i = 1, nc_b
             box%cc(i, j, k, i_rhs) = gauss_lpl(gs, a$D_r_cc(box, [i, j, k]))
          end do
       end do
    end do
    if (lsf_mode > 0) then
       do k = 0, nc_b+1
          do j = 0, nc_b+1
             do i = 0, nc_b+1
                r = a$D_r_cc(box, [i, j, k])
                box%cc(i, j, k, i_lsf) = norm2(r - lsf_c) - lsf_rad
                box%cc(i, j, k, i_bval) = 1.0_dp + r(1)
             end do
          end do
       end do
    end if
    if (use_eps) then
       do k = 0, nc_b+1
          do j = 0, nc_b+1
             do i = 0, nc_b+1
                r = a$D_r_cc(box, [i, j, k])
                if (r($D) < eps_cut) then
                   box%cc(i, j, k, i_eps) = 1.0_dp
                else
                   box%cc(i, j, k, i_eps) = eps_hi
                end if
             end do
          end do
       end do
    end if
#endif
  end subroutine set_init

  !> Dump format (little-endian stream):
  !>   char(4) 'AFGT'; int32 version=1, dim, nc, n_var, min_lvl, highest_lvl,
  !>   n_boxes; real64 dr_base;
  !>   n_boxes records of int32 [id, lvl, ix(dim), parent, children(2^dim),
  !>   neighbors(2*dim)] in level order (ids order within a level);
  !>   then for iv = 1..n_var, for each box in the same order, the whole
  !>   cc(0:nc_l+1, ..., iv) array in Fortran (column-major) order.
  subroutine write_dump(fname)
    character(len=*), intent(in) :: fname
    integer :: lu, lvl, n, id, n_boxes, iv, min_lvl

    min_lvl = lbound(tree%lvls, 1)
    n_boxes = 0
    do lvl = min_lvl, tree%highest_lvl
       n_boxes = n_boxes + size(tree%lvls(lvl)%ids)
    end do

    open(newunit=lu, file=fname, access='stream', form='unformatted', &
         status='replace', action='write')
    write(lu) 'AFGT'
    write(lu) 1, $D, tree%n_cell, n_var, min_lvl, tree%highest_lvl, n_boxes
    write(lu) tree%dr_base
    do lvl = min_lvl, tree%highest_lvl
       do n = 1, size(tree%lvls(lvl)%ids)
          id = tree%lvls(lvl)%ids(n)
          write(lu) id, tree%boxes(id)%lvl, tree%boxes(id)%ix, &
               tree%boxes(id)%parent, tree%boxes(id)%children, &
               tree%boxes(id)%neighbors
       end do
    end do
    do iv = 1, n_var
       do lvl = min_lvl, tree%highest_lvl
          do n = 1, size(tree%lvls(lvl)%ids)
             id = tree%lvls(lvl)%ids(n)
#if $D == 2
             write(lu) tree%boxes(id)%cc(:, :, iv)
#elif $D == 3
             write(lu) tree%boxes(id)%cc(:, :, :, iv)
#endif
          end do
       end do
    end do
    close(lu)
  end subroutine write_dump

end program ref_harness_$Dd
