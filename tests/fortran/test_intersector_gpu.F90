! Drives the Fortran interface the way the reference's serial tests do
! (src/tests/test_tri_intersector.F90:84-104, test_tet_intersector.F90:84-104,
! test_element_intersector.F90, test_interval_intersector.F90).
! Usage: test_intersector_gpu mesh_a.bin mesh_b.bin [mode]
!   mode "pairs" (default): tree finder, then intersect_tris / intersect_tets
!     per pair, summing triangle_area / tetrahedron_volume of the pieces;
!   mode "callback": the supermesh callback driver, summing in the callback
!     (any element types);
!   mode "badnode": the callback driver on mesh A with an out-of-range node
!     id, which must leave an error message and make no callback;
!   mode "mixed": simplex mesh A against quad / hex mesh B (R-tree finder,
!     intersect_tri_quad / intersect_tet_hex), or 1D meshes (sort finder,
!     intersect_intervals);
!   mode "elements": every closed-overlap pair through intersect_elements
!     (any element types), summing simplex_volume of the pieces;
!   mode "tree_api": one-element queries through the stateful tree finder
!     (set_input / find / query_output / get_output / reset) and an
!     rtree_type (allocate / query / deallocate), checked against the CSR
!     finders, plus the ragged brute force form; prints the pair count;
!   mode "front" / "finder": advancing_front_intersection_finder /
!     intersection_finder only; the third number printed is then a checksum
!     of the CSR, sum of indices(i) * i plus sum of indptr(i) * (i + 7).
! Mesh file (stream, little endian): int32 nnodes, dim, nelements, loc;
! real(8) positions(dim, nnodes); int32 enlist(loc, nelements) (1-based).
! Prints: dim, number of candidate pairs, number of pieces, measure sum.
module test_intersector_state
  ! State shared with the callback: a module procedure, not an internal one,
  ! so that passing it needs no trampoline on the stack.
  use libsupermesh
  implicit none
  real(kind = real_kind), dimension(:, :), allocatable :: pa, pb
  integer, dimension(:, :), allocatable :: ea, eb
  real(kind = real_kind) :: total
  integer :: npieces, npairs, dim

contains

  subroutine accumulate(element_a, element_b, elements_c, nodes_b, ele_a, ele_b, local)
    real(kind = real_kind), dimension(:, :), intent(in) :: element_a, element_b
    real(kind = real_kind), dimension(:, :, :), intent(in) :: elements_c
    integer, dimension(:), intent(in) :: nodes_b
    integer, intent(in) :: ele_a, ele_b
    logical, intent(in) :: local
    integer :: k
    if(any(nodes_b /= eb(:, ele_b)) .or. any(element_a /= pa(:, ea(:, ele_a))) .or. &
      & any(element_b /= pb(:, eb(:, ele_b))) .or. .not. local) then
      print "(a)", "error: callback arguments disagree with the meshes"
      stop 1
    end if
    npairs = npairs + 1
    npieces = npieces + size(elements_c, 3)
    do k = 1, size(elements_c, 3)
      if(dim == 2) then
        total = total + triangle_area(elements_c(:, :, k))
      else
        total = total + tetrahedron_volume(elements_c(:, :, k))
      end if
    end do
  end subroutine accumulate

end module test_intersector_state

program test_intersector_gpu
  use libsupermesh
  use test_intersector_state
  implicit none
  character(len = 512) :: fa, fb, mode
  integer, dimension(:), allocatable :: indices, indptr
  real(kind = real_kind), dimension(2, 3, tri_buf_size) :: tris_c
  real(kind = real_kind), dimension(3, 4, tet_buf_size) :: tets_c
  real(kind = real_kind), dimension(2, 3, 46) :: tris_q
  real(kind = real_kind), dimension(:, :, :), allocatable :: tets_h
  real(kind = real_kind), dimension(2) :: interval_c
  integer :: a, i, b, c, n
  integer(kind = 8) :: chk
  integer, dimension(:), allocatable :: indices2, indptr2, eles
  type(intersections), dimension(:), allocatable :: map_ab
  type(rtree_type) :: rt

  call get_command_argument(1, fa)
  call get_command_argument(2, fb)
  mode = "pairs"
  if(command_argument_count() >= 3) call get_command_argument(3, mode)
  call read_mesh(trim(fa), pa, ea)
  call read_mesh(trim(fb), pb, eb)
  dim = size(pa, 1)
  total = 0.0_real_kind
  npieces = 0
  npairs = 0

  select case(trim(mode))
  case("callback")
    call supermesh(pa, ea, pb, eb, accumulate)
    call check_error()
  case("badnode")
    ! an out-of-range node id must fail with a message and no callbacks
    ea(1, 1) = size(pa, 2) + 5
    call supermesh(pa, ea, pb, eb, accumulate)
    if(len(libsupermesh_gpu_error()) == 0 .or. npairs /= 0) then
      print "(a)", "error: supermesh accepted an invalid node id"
      stop 1
    end if
  case("mixed")
    allocate(indptr(size(ea, 2) + 1))
    if(dim == 1) then
      call sort_intersection_finder(pa, ea, pb, eb, indices, indptr)
    else
      call rtree_intersection_finder(pa, ea, pb, eb, indices, indptr)
    end if
    call check_error()
    allocate(tets_h(3, 4, 729))
    do a = 1, size(ea, 2)
      do i = indptr(a), indptr(a + 1) - 1
        b = indices(i)
        npairs = npairs + 1
        if(dim == 1) then
          call intersect_intervals(pa(1, ea(:, a)), pb(1, eb(:, b)), interval_c, n)
          if(n > 0) total = total + interval_size(interval_c)
        else if(dim == 2) then
          call intersect_tri_quad(pa(:, ea(:, a)), pb(:, eb(:, b)), tris_q, n)
          do c = 1, n
            total = total + triangle_area(tris_q(:, :, c))
          end do
        else
          call intersect_tet_hex(pa(:, ea(:, a)), pb(:, eb(:, b)), tets_h, n)
          do c = 1, n
            total = total + tetrahedron_volume(tets_h(:, :, c))
          end do
        end if
        npieces = npieces + n
      end do
    end do
    call check_error()
  case("elements")
    allocate(indptr(size(ea, 2) + 1))
    call intersection_finder(pa, ea, pb, eb, indices, indptr)
    call check_error()
    allocate(tets_h(dim, dim + 1, max_n_simplices_c(dim, size(ea, 1), size(eb, 1))))
    do a = 1, size(ea, 2)
      do i = indptr(a), indptr(a + 1) - 1
        b = indices(i)
        npairs = npairs + 1
        call intersect_elements(pa(:, ea(:, a)), pb(:, eb(:, b)), tets_h, n)
        do c = 1, n
          total = total + simplex_volume(tets_h(:, :, c))
        end do
        npieces = npieces + n
      end do
    end do
    call check_error()
  case("tree_api")
    allocate(indptr(size(ea, 2) + 1), indptr2(size(ea, 2) + 1), map_ab(size(ea, 2)))
    call tree_intersection_finder(pa, ea, pb, eb, indices, indptr)
    call rtree_intersection_finder(pa, ea, pb, eb, indices2, indptr2)
    call brute_force_intersection_finder(pa, ea, pb, eb, map_ab)
    call check_error()
    call tree_intersection_finder_set_input(pb, eb)
    call allocate(rt, pb, eb)
    call check_error()
    do a = 1, size(ea, 2)
      call tree_intersection_finder_find(pa(:, ea(:, a)))
      call tree_intersection_finder_query_output(n)
      if(n /= indptr(a + 1) - indptr(a) .or. map_ab(a)%n /= n) then
        print "(a,i0)", "error: tree row length, element ", a
        stop 1
      end if
      do i = 1, n
        call tree_intersection_finder_get_output(b, i)
        if(b /= indices(indptr(a) + i - 1) .or. map_ab(a)%v(i) /= b) then
          print "(a,i0)", "error: tree row entry, element ", a
          stop 1
        end if
      end do
      call query(rt, pa(:, ea(:, a)), eles)
      if(size(eles) /= indptr2(a + 1) - indptr2(a)) then
        print "(a,i0)", "error: rtree row length, element ", a
        stop 1
      end if
      if(any(eles /= indices2(indptr2(a):indptr2(a + 1) - 1))) then
        print "(a,i0)", "error: rtree row entry, element ", a
        stop 1
      end if
      npairs = npairs + n
      npieces = npieces + size(eles)
    end do
    call check_error()
    call tree_intersection_finder_reset()
    call deallocate(rt)
    call deallocate(map_ab)
  case("front", "finder")
    allocate(indptr(size(ea, 2) + 1))
    if(trim(mode) == "front") then
      call advancing_front_intersection_finder(pa, ea, pb, eb, indices, indptr)
    else
      call intersection_finder(pa, ea, pb, eb, indices, indptr)
    end if
    call check_error()
    chk = 0
    do i = 1, size(indices)
      chk = chk + int(indices(i), 8) * int(i, 8)
    end do
    do i = 1, size(indptr)
      chk = chk + int(indptr(i), 8) * int(i + 7, 8)
    end do
    print "(i0,1x,i0,1x,i0,1x,es25.17)", dim, size(indices), chk, 0.0_real_kind
    stop
  case default
    allocate(indptr(size(ea, 2) + 1))
    call tree_intersection_finder(pa, ea, pb, eb, indices, indptr)
    call check_error()
    do a = 1, size(ea, 2)
      do i = indptr(a), indptr(a + 1) - 1
        b = indices(i)
        npairs = npairs + 1
        if(dim == 2) then
          call intersect_tris(pa(:, ea(:, a)), pb(:, eb(:, b)), tris_c, n)
          do c = 1, n
            total = total + triangle_area(tris_c(:, :, c))
          end do
        else
          call intersect_tets(pa(:, ea(:, a)), pb(:, eb(:, b)), tets_c, n)
          do c = 1, n
            total = total + tetrahedron_volume(tets_c(:, :, c))
          end do
        end if
        npieces = npieces + n
      end do
    end do
  end select
  print "(i0,1x,i0,1x,i0,1x,es25.17)", dim, npairs, npieces, total

contains


  subroutine check_error()
    if(len(libsupermesh_gpu_error()) > 0) then
      print "(a,a)", "error: ", libsupermesh_gpu_error()
      stop 1
    end if
  end subroutine check_error

  subroutine read_mesh(path, pos, enl)
    character(len = *), intent(in) :: path
    real(kind = real_kind), dimension(:, :), allocatable, intent(out) :: pos
    integer, dimension(:, :), allocatable, intent(out) :: enl
    integer :: u, nn, d, ne, loc
    open(newunit = u, file = path, access = "stream", form = "unformatted", status = "old")
    read(u) nn, d, ne, loc
    allocate(pos(d, nn), enl(loc, ne))
    read(u) pos
    read(u) enl
    close(u)
  end subroutine read_mesh

end program test_intersector_gpu
