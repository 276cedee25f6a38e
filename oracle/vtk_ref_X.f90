! Test program: the VTK output of a tree written by the reference's own XML
! writer module m_vtk (src/m_vtk.f90, compiled from the reference sources by
! oracle/Makefile), with the arrays assembled as a$D_write_vtk assembles them
! (src/m_aX_io.f90:115-310; that module itself cannot be built here, as it
! also holds the Silo writer, whose library is absent).  Test infrastructure:
! the checker of afivo_amd/vtk.py's write_vtu, through the fixtures that
! tests/golden/make_golden.py writes with it.
!
! Tree: boxes of 4^$D cells on [r0, r0 + 1]^$D (r0 from the command line),
! coarse grids down to 2 cells, uniform to level 2 and refined to level 4 (2D)
! or 3 (3D) where r_min < r0 + 0.5 in every coordinate; three cell variables "phi",
! "rhs", "tmp" and one face variable "flux" with deterministic values (ghost
! cells included).  Writes <out>.tree (the tree/state format of
! afivo_amd/treefile.py), <out>.fc (the face variable: per box in record
! order fx, fy [, fz] of box%fx(:, :, [:,] 1) ..., Fortran order) and
! <out>.vtu, with n_cycle 7 and time 0.125, all cell variables and the face
! variable.  Usage: vtk_ref_$Dd <out> <r0> [cyl]
program vtk_ref
  use m_a$D_t
  use m_a$D_core
  use m_a$D_utils
  use m_vtk, only: vtk_t, vtk_ini_xml, vtk_dat_xml, vtk_unstr_geo_xml, vtk_unstr_con_xml, &
       vtk_var_r8_xml, vtk_unstr_geo_xml_close, vtk_end_xml

  implicit none

  integer, parameter :: nc = 4
#if $D == 2
  integer, parameter :: l_max = 4
#elif $D == 3
  integer, parameter :: l_max = 3
#endif
  type(a$D_t)        :: tree
  type(ref_info_t)   :: ref_info
  integer            :: ix_list($D, 1), nb_list(2*$D, 1), coord, d
  real(dp)           :: r0
  character(len=256) :: out, arg

  call get_command_argument(1, out)
  call get_command_argument(2, arg)
  read(arg, *) r0
  coord = a5_xyz
  if (command_argument_count() > 2) then
     call get_command_argument(3, arg)
     if (arg == "cyl") coord = a5_cyl
  end if

  call a$D_init(tree, nc, 3, 1, 1.0_dp / nc, r_min=[(r0, d = 1, $D)], &
       coarsen_to=2, coord=coord, cc_names=["phi", "rhs", "tmp"], fc_names=["flux"])
  ix_list(:, 1) = 1
  nb_list(:, 1) = -1
  call a$D_set_base(tree, ix_list, nb_list)
  do
     call a$D_adjust_refinement(tree, ref_routine, ref_info)
     if (ref_info%n_add == 0) exit
  end do
  call a$D_loop_box(tree, set_values)

  call write_tree(trim(out) // ".tree")
  call write_faces(trim(out) // ".fc")
  call write_vtk(trim(out) // ".vtu", 7, 0.125_dp, [1, 2, 3], [1])

contains

  subroutine ref_routine(boxes, id, ref_flag)
    type(box$D_t), intent(in) :: boxes(:)
    integer, intent(in)       :: id
    integer, intent(inout)    :: ref_flag

    if (boxes(id)%lvl < 2) then
       ref_flag = a5_do_ref
    else if (boxes(id)%lvl < l_max .and. all(boxes(id)%r_min < r0 + 0.5_dp)) then
       ref_flag = a5_do_ref
    end if
  end subroutine ref_routine

  real(dp) function wave(box, i, j, k, iv)
    type(box$D_t), intent(in) :: box
    integer, intent(in)       :: i, j, k, iv
    integer                   :: s
    s = 1000 * box%lvl + 37 * box%ix(1) + 101 * box%ix(2)
#if $D == 3
    s = s + 211 * box%ix(3)
#endif
    wave = sin(0.713_dp * s + 1.37_dp * i + 2.11_dp * j + 0.53_dp * k + 3.7_dp * iv)
  end function wave

  subroutine set_values(box)
    type(box$D_t), intent(inout) :: box
    integer                      :: i, j, k, iv, n

    n = box%n_cell
    k = 0
#if $D == 2
    do j = 0, n + 1
       do i = 0, n + 1
          do iv = 1, 3
             box%cc(i, j, iv) = wave(box, i, j, k, iv)
          end do
       end do
    end do
    do j = 1, n
       do i = 1, n + 1
          box%fx(i, j, 1) = wave(box, i, j, k, 11)
          box%fy(j, i, 1) = wave(box, j, i, k, 12)
       end do
    end do
#elif $D == 3
    do k = 0, n + 1
       do j = 0, n + 1
          do i = 0, n + 1
             do iv = 1, 3
                box%cc(i, j, k, iv) = wave(box, i, j, k, iv)
             end do
          end do
       end do
    end do
    do k = 1, n
       do j = 1, n
          do i = 1, n + 1
             box%fx(i, j, k, 1) = wave(box, i, j, k, 11)
             box%fy(j, i, k, 1) = wave(box, j, i, k, 12)
             box%fz(j, k, i, 1) = wave(box, j, k, i, 13)
          end do
       end do
    end do
#endif
  end subroutine set_values

  !> the tree/state file of afivo_amd/treefile.py: header, records and cell
  !> data in level order
  subroutine write_tree(fname)
    character(len=*), intent(in) :: fname
    integer :: lu, lvl, n, id, n_boxes, iv, min_lvl

    min_lvl = lbound(tree%lvls, 1)
    n_boxes = 0
    do lvl = min_lvl, tree%highest_lvl
       n_boxes = n_boxes + size(tree%lvls(lvl)%ids)
    end do
    open(newunit=lu, file=fname, access='stream', form='unformatted', status='replace')
    write(lu) 'AFGT'
    write(lu) 1, $D, tree%n_cell, 3, min_lvl, tree%highest_lvl, n_boxes
    write(lu) tree%dr_base
    do lvl = min_lvl, tree%highest_lvl
       do n = 1, size(tree%lvls(lvl)%ids)
          id = tree%lvls(lvl)%ids(n)
          write(lu) id, tree%boxes(id)%lvl, tree%boxes(id)%ix, tree%boxes(id)%parent, &
               tree%boxes(id)%children, tree%boxes(id)%neighbors
       end do
    end do
    do iv = 1, 3
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
  end subroutine write_tree

  subroutine write_faces(fname)
    character(len=*), intent(in) :: fname
    integer :: lu, lvl, n, id

    open(newunit=lu, file=fname, access='stream', form='unformatted', status='replace')
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

  !> The leaves of levels 1..highest, one pixel / voxel per cell with its own
  !> corner nodes, the cell variables icc and the face variables ifc
  !> averaged to the cell centre, in the calls to m_vtk of a$D_write_vtk
  subroutine write_vtk(fname, n_cycle, time, icc, ifc)
    character(len=*), intent(in) :: fname
    integer, intent(in)          :: n_cycle, icc(:), ifc(:)
    real(dp), intent(in)         :: time
    integer, parameter           :: n_ch = 2**$D
    integer                      :: lvl, bc, bn, n, n_cells, n_nodes, ig, i, j, k, id
    integer                      :: n_ix, c_ix, n_grids, cell_ix, node_ix, n_cc, n_fc, m
    real(dp), allocatable        :: coords(:), cc_vars(:, :)
    integer, allocatable         :: offsets(:), connects(:), cell_types(:)
    character(len=100), allocatable :: var_names(:)
    type(vtk_t)                  :: vtkf
    character(len=2)             :: sfx($D)

    n_cc = size(icc)
    n_fc = size(ifc)
    allocate(var_names(n_cc + n_fc * $D))
    var_names(1:n_cc) = tree%cc_names(icc)
#if $D == 2
    sfx = ["_x", "_y"]
    if (tree%coord_t == a5_cyl) sfx = ["_r", "_z"]
#elif $D == 3
    sfx = ["_x", "_y", "_z"]
#endif
    do i = 1, n_fc
       do m = 1, $D
          var_names(n_cc + (i-1) * $D + m) = trim(tree%fc_names(ifc(i))) // sfx(m)
       end do
    end do

    bc = tree%n_cell
    bn = tree%n_cell + 1
    n_grids = 0
    do lvl = 1, tree%highest_lvl
       n_grids = n_grids + size(tree%lvls(lvl)%leaves)
    end do
    n_nodes = bn**$D * n_grids
    n_cells = bc**$D * n_grids
    allocate(coords($D * n_nodes), cc_vars(n_cells, n_cc + n_fc * $D))
    allocate(offsets(n_cells), cell_types(n_cells), connects(n_ch * n_cells))
    cell_types = merge(8, 11, $D == 2)

    ig = 0
    do lvl = 1, tree%highest_lvl
       do n = 1, size(tree%lvls(lvl)%leaves)
          id = tree%lvls(lvl)%leaves(n)
          ig = ig + 1
          cell_ix = (ig - 1) * bc**$D
          node_ix = (ig - 1) * bn**$D
#if $D == 2
          k = 1
          do j = 1, bn
             do i = 1, bn
                n_ix = 2 * (node_ix + (j-1) * bn + i)
                coords(n_ix-1:n_ix) = tree%boxes(id)%r_min + [i-1, j-1] * tree%boxes(id)%dr
             end do
          end do
          do j = 1, bc
             do i = 1, bc
                n_ix = node_ix + (j-1) * bn + i - 1
                c_ix = cell_ix + (j-1) * bc + i
                cc_vars(c_ix, 1:n_cc) = tree%boxes(id)%cc(i, j, icc)
                cc_vars(c_ix, n_cc+1::$D) = 0.5_dp * sum(tree%boxes(id)%fx(i:i+1, j, ifc))
                cc_vars(c_ix, n_cc+2::$D) = 0.5_dp * sum(tree%boxes(id)%fy(i, j:j+1, ifc))
                offsets(c_ix) = n_ch * c_ix
                connects(n_ch*(c_ix-1)+1:n_ch*c_ix) = [n_ix, n_ix+1, n_ix+bn, n_ix+bn+1]
             end do
          end do
#elif $D == 3
          do k = 1, bn
             do j = 1, bn
                do i = 1, bn
                   n_ix = 3 * (node_ix + (k-1) * bn**2 + (j-1) * bn + i)
                   coords(n_ix-2:n_ix) = tree%boxes(id)%r_min + [i-1, j-1, k-1] * tree%boxes(id)%dr
                end do
             end do
          end do
          do k = 1, bc
             do j = 1, bc
                do i = 1, bc
                   n_ix = node_ix + (k-1) * bn**2 + (j-1) * bn + i - 1
                   c_ix = cell_ix + (k-1) * bc**2 + (j-1) * bc + i
                   cc_vars(c_ix, 1:n_cc) = tree%boxes(id)%cc(i, j, k, icc)
                   cc_vars(c_ix, n_cc+1::$D) = 0.5_dp * sum(tree%boxes(id)%fx(i:i+1, j, k, ifc))
                   cc_vars(c_ix, n_cc+2::$D) = 0.5_dp * sum(tree%boxes(id)%fy(i, j:j+1, k, ifc))
                   cc_vars(c_ix, n_cc+3::$D) = 0.5_dp * sum(tree%boxes(id)%fz(i, j, k:k+1, ifc))
                   offsets(c_ix) = n_ch * c_ix
                   connects(n_ch*(c_ix-1)+1:n_ch*c_ix) = [n_ix, n_ix+1, n_ix+bn, n_ix+bn+1, &
                        n_ix+bn**2, n_ix+bn**2+1, n_ix+bn**2+bn, n_ix+bn**2+bn+1]
                end do
             end do
          end do
#endif
       end do
    end do

    call vtk_ini_xml(vtkf, fname, 'UnstructuredGrid')
    call vtk_dat_xml(vtkf, "UnstructuredGrid", .true.)
    call vtk_unstr_geo_xml(vtkf, coords, n_nodes, n_cells, $D, n_cycle, time)
    call vtk_unstr_con_xml(vtkf, connects, offsets, cell_types, n_cells)
    call vtk_dat_xml(vtkf, "CellData", .true.)
    do n = 1, n_cc + n_fc * $D
       call vtk_var_r8_xml(vtkf, trim(var_names(n)), cc_vars(:, n), n_cells)
    end do
    call vtk_dat_xml(vtkf, "CellData", .false.)
    call vtk_unstr_geo_xml_close(vtkf)
    call vtk_dat_xml(vtkf, "UnstructuredGrid", .false.)
    call vtk_end_xml(vtkf)
  end subroutine write_vtk
end program vtk_ref
