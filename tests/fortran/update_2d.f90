! Test program of the drop-in's tree update (afivo_amd/fortran/m_aX_mg.f90):
! a device-resident solve, then one refinement step done twice from the
! same data: on a host copy as afivo's programs do it (a$D_adjust_refinement,
! then a$D_prolong1 of phi into every added box, lowest level first), and on
! the device with mg2_device_update.  Test infrastructure.
!
! Set-up as the reference examples (examples/poisson_basic_2d.f90:41-98):
! unit square, 8x8 boxes, uniform to level 3, refined to level 5 where
! dr^2 max|rhs| > 1e-3; rhs = gauss_lpl of two Gaussians; Dirichlet zero.
! The step refines to level 6 where dr^2 max|rhs| > 1e-3 and derefines
! where it is below 1e-4.
! Prints: boxes before and after, added boxes, kept boxes whose phi differs
! in any bit (ghost cells included), added boxes whose phi interior differs
! in any bit, and the relative residual after FMG (have_guess) and two
! V-cycles on the new tree.
program update_2d
  use m_a2_t
  use m_a2_core
  use m_a2_mg
  use m_a2_utils
  use m_a2_gc
  use m_a2_prolong
  use m_gaussians

  implicit none

  integer, parameter :: i_phi = 1, i_rhs = 2, i_tmp = 3
  type(a2_t)       :: tree, host
  type(ref_info_t) :: ref_info, ref_host
  type(mg2_t)      :: mg
  type(gauss_t)    :: gs
  integer          :: ix_list(2, 1), nb_list(4, 1), lvl, i, id, n_before, n_kept_bad, n_new_bad
  integer          :: n_added, highest_prev
  real(dp)         :: rmin, rmax, rhs_max, res
  logical          :: growing

  call gauss_init(gs, [1.0_dp, 1.0_dp], [0.04_dp, 0.04_dp], &
       reshape([0.25_dp, 0.25_dp, 0.75_dp, 0.75_dp], [2, 2]))
  call a2_init(tree, 8, 3, 0, 1.0_dp / 8, coarsen_to=2)
  ix_list(:, 1) = 1
  nb_list(:, 1) = -1
  call a2_set_base(tree, ix_list, nb_list)
  growing = .true.
  do
     call a2_adjust_refinement(tree, ref_routine, ref_info)
     if (ref_info%n_add == 0) exit
  end do
  call a2_loop_box(tree, set_init)

  mg%i_phi = i_phi
  mg%i_rhs = i_rhs
  mg%i_tmp = i_tmp
  mg%sides_bc => a2_bc_dirichlet_zero
  mg%n_cycle_base = 50
  mg%device_resident = .true.
  call mg2_init_mg(mg)

  call mg2_fas_fmg(tree, mg, .true., .false.)
  call mg2_to_host(tree, mg)
  n_before = a2_num_boxes_used(tree)
  highest_prev = tree%highest_id

  ! on the host: the data is moved by afivo itself
  host = tree
  growing = .false.
  call a2_adjust_refinement(host, ref_routine, ref_host)
  do lvl = lbound(ref_host%lvls, 1), ubound(ref_host%lvls, 1)
     if (.not. allocated(ref_host%lvls(lvl)%add)) cycle
     do i = 1, size(ref_host%lvls(lvl)%add)
        id = ref_host%lvls(lvl)%add(i)
        call a2_prolong1(host%boxes(host%boxes(id)%parent), host%boxes(id), i_phi)
     end do
  end do

  ! on the device
  call a2_adjust_refinement(tree, ref_routine, ref_info)
  call mg2_device_update(tree, 1, [i_phi])
  call mg2_to_host(tree, mg)

  n_kept_bad = 0
  n_new_bad = 0
  n_added = 0
  do lvl = 1, tree%highest_lvl
     do i = 1, size(tree%lvls(lvl)%ids)
        id = tree%lvls(lvl)%ids(i)
        if (id > highest_prev) then
           n_added = n_added + 1
           if (any(transfer(tree%boxes(id)%cc(1:8, 1:8, i_phi), 1_8, 64) /= &
                transfer(host%boxes(id)%cc(1:8, 1:8, i_phi), 1_8, 64))) n_new_bad = n_new_bad + 1
        else
           if (any(transfer(tree%boxes(id)%cc(:, :, i_phi), 1_8, 100) /= &
                transfer(host%boxes(id)%cc(:, :, i_phi), 1_8, 100))) n_kept_bad = n_kept_bad + 1
        end if
     end do
  end do

  ! continue on the new tree: rhs of every box from the host, phi from the device
  call a2_loop_box(tree, set_rhs)
  call mg2_to_device(tree, i_rhs)
  call a2_tree_min_cc(tree, i_rhs, rmin)
  call a2_tree_max_cc(tree, i_rhs, rmax)
  rhs_max = max(abs(rmin), abs(rmax))
  call mg2_fas_fmg(tree, mg, .true., .true.)
  call mg2_fas_vcycle(tree, mg, tree%highest_lvl, .true.)
  call mg2_fas_vcycle(tree, mg, tree%highest_lvl, .true.)
  call mg2_to_host(tree, mg)
  call a2_tree_min_cc(tree, i_tmp, rmin)
  call a2_tree_max_cc(tree, i_tmp, rmax)
  res = max(abs(rmin), abs(rmax)) / rhs_max

  write(*, '(A, I0, A, I0, A, I0, A, I0, A, I0, A, ES12.4)') 'before ', n_before, &
       ' after ', a2_num_boxes_used(tree), ' added ', n_added, ' kept_bad ', n_kept_bad, &
       ' new_bad ', n_new_bad, ' rel_res ', res
  call mg2_device_release()

contains

  subroutine ref_routine(boxes, id, ref_flag)
    type(box2_t), intent(in) :: boxes(:)
    integer, intent(in)      :: id
    integer, intent(inout)   :: ref_flag
    integer                  :: i, j
    real(dp)                 :: rmax

    if (boxes(id)%lvl < 3) then
       ref_flag = a5_do_ref
       return
    end if
    rmax = 0
    do j = 1, 8
       do i = 1, 8
          rmax = max(rmax, abs(gauss_lpl(gs, a2_r_cc(boxes(id), [i, j]))))
       end do
    end do
    if (growing) then
       if (boxes(id)%lvl < 5 .and. boxes(id)%dr**2 * rmax > 1.0e-3_dp) ref_flag = a5_do_ref
    else
       if (boxes(id)%lvl < 6 .and. boxes(id)%dr**2 * rmax > 1.0e-3_dp) then
          ref_flag = a5_do_ref
       else if (boxes(id)%dr**2 * rmax < 1.0e-4_dp) then
          ref_flag = a5_rm_ref
       end if
    end if
  end subroutine ref_routine

  subroutine set_init(box)
    type(box2_t), intent(inout) :: box
    box%cc = 0.0_dp
    call set_rhs(box)
  end subroutine set_init

  subroutine set_rhs(box)
    type(box2_t), intent(inout) :: box
    integer                     :: i, j

    do j = 1, box%n_cell
       do i = 1, box%n_cell
          box%cc(i, j, i_rhs) = gauss_lpl(gs, a2_r_cc(box, [i, j]))
       end do
    end do
  end subroutine set_rhs
end program update_2d
