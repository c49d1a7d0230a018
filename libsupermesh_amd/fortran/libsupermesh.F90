! libsupermesh.F90 -- Fortran interface of the MI355X supermeshing engine.
!
! Mirrors the reference's `use libsupermesh` surface for the hot path
! (src/libsupermesh.F90:24-43, src/Supermesh.F90:72-81,
! src/Intersection_Finder.F90:117-131) on top of the C ABI
! (include/libsupermesh-c.h) through ISO_C_BINDING.  Arrays keep the
! reference's Fortran layouts: positions(dim, nnodes), enlist(loc, nelements)
! with 1-based node ids, tri(2, 3), tet(3, 4), and CSR outputs with 1-based
! indices and pointers, as the reference's *_csr_sparsity routines return.
! Errors: the reference aborts; here the routines return with empty results
! and libsupermesh_gpu_error() holds the message.
module libsupermesh

  use iso_c_binding, only : c_double, c_int, c_long, c_char, c_ptr, c_null_char, &
    & c_f_pointer, c_associated, c_loc, c_funloc, c_funptr, c_null_ptr, c_size_t

  implicit none

  private

  integer, parameter, public :: real_kind = c_double
  integer, parameter, public :: tri_buf_size = 22
  integer, parameter, public :: tet_buf_size = 81

  integer, parameter, public :: interval_buf_size = 1

  public :: tree_intersection_finder, rtree_intersection_finder, &
    & sort_intersection_finder, advancing_front_intersection_finder, &
    & intersection_finder, brute_force_intersection_finder, intersect_tris, intersect_tets, &
    & triangle_area, tetrahedron_volume, intersect_intervals, interval_size, &
    & intersect_tri_quad, intersect_tet_hex, supermesh, libsupermesh_gpu_error
  public :: intersections, tree_type, rtree_type, allocate, query, deallocate
  public :: intersect_quads, intersect_hexes, intersect_simplices, intersect_elements, &
    & simplex_volume, max_n_tris_c, max_n_tets_c, max_n_simplices_c

  interface max_n_tris_c
    module procedure max_n_tris_c_tri, max_n_tris_c_poly
  end interface max_n_tris_c

  interface max_n_simplices_c
    module procedure max_n_simplices_c_simplices, max_n_simplices_c_elements
  end interface max_n_simplices_c
  public :: tree_intersection_finder_set_input, tree_intersection_finder_find, &
    & tree_intersection_finder_query_output, tree_intersection_finder_get_output, &
    & tree_intersection_finder_reset
  public :: rtree_intersection_finder_set_input, rtree_intersection_finder_find, &
    & rtree_intersection_finder_query_output, rtree_intersection_finder_get_output, &
    & rtree_intersection_finder_reset

  ! Ragged candidate lists, one per A element (Intersections.F90:35-40).
  type intersections
    integer, dimension(:), pointer :: v => null()
    integer :: n = 0
  end type intersections

  ! A search hierarchy over one mesh, kept on the device
  ! (libsupermesh_gpu_tree; Intersection_Finder.F90:164-172).
  type tree_type
    integer :: dim = 0
    type(c_ptr) :: handle = c_null_ptr
  end type tree_type

  type rtree_type
    type(c_ptr) :: handle = c_null_ptr
  end type rtree_type

  interface allocate
    module procedure allocate_tree, allocate_rtree
  end interface allocate

  interface query
    module procedure query_tree_allocatable, query_tree_pointer, query_rtree_allocatable, &
      & query_rtree_pointer
  end interface query

  interface deallocate
    module procedure deallocate_tree, deallocate_rtree, deallocate_intersections, &
      & deallocate_intersections_rank_1
  end interface deallocate

  ! state of the set_input / find / query_output / get_output / reset calls
  type(tree_type), save :: state_tree
  type(rtree_type), save :: state_rtree
  integer, dimension(:), allocatable, save :: state_tree_eles, state_rtree_eles

  ! The per-pair callback of supermesh: the interface of the reference's
  ! intersection_calculation (Parallel_Supermesh.F90:97-113).
  abstract interface
    subroutine intersection_calculation_interface(element_a, element_b, elements_c, nodes_b, &
        & ele_a, ele_b, local)
      import :: c_double
      real(kind = c_double), dimension(:, :), intent(in) :: element_a, element_b
      real(kind = c_double), dimension(:, :, :), intent(in) :: elements_c
      integer, dimension(:), intent(in) :: nodes_b
      integer, intent(in) :: ele_a, ele_b
      logical, intent(in) :: local
    end subroutine intersection_calculation_interface
  end interface

  ! libsupermesh_gpu_mesh (include/libsupermesh_gpu.h)
  type, bind(c) :: gpu_mesh
    integer(kind = c_int) :: dim
    integer(kind = c_long) :: nnodes, nelements
    type(c_ptr) :: positions, enlist
  end type gpu_mesh

  procedure(intersection_calculation_interface), pointer, save :: user_calculation => null()
  integer, save :: callback_dim = 0, callback_loc_a = 0, callback_loc_b = 0
  type(c_ptr), save :: gpu_context = c_null_ptr

  interface tree_intersection_finder
    module procedure tree_intersection_finder_csr_sparsity, tree_intersection_finder_intersections
  end interface tree_intersection_finder

  interface rtree_intersection_finder
    module procedure rtree_intersection_finder_csr_sparsity, rtree_intersection_finder_intersections
  end interface rtree_intersection_finder

  interface brute_force_intersection_finder
    module procedure brute_force_intersection_finder_csr_sparsity, &
      & brute_force_intersection_finder_intersections
  end interface brute_force_intersection_finder

  interface intersect_tris
    module procedure intersect_tris_real
  end interface intersect_tris

  interface intersect_tets
    module procedure intersect_tets_real
  end interface intersect_tets

  interface sort_intersection_finder
    module procedure sort_intersection_finder_csr_sparsity, sort_intersection_finder_intersections
  end interface sort_intersection_finder

  interface advancing_front_intersection_finder
    module procedure advancing_front_intersection_finder_csr_sparsity, &
      & advancing_front_intersection_finder_intersections
  end interface advancing_front_intersection_finder

  interface intersection_finder
    module procedure intersection_finder_csr_sparsity, intersection_finder_intersections
  end interface intersection_finder

  ! finders behind finder_csr
  integer, parameter :: finder_tree = 0, finder_rtree = 1, finder_sort = 2, finder_front = 3

  interface
    subroutine c_tree_set_input(nnodes_a, dim_a, nelements_a, loc_a, nnodes_b, dim_b, &
        & nelements_b, loc_b, positions_a, enlist_a, positions_b, enlist_b) &
        & bind(c, name = "libsupermesh_tree_intersection_finder_set_input")
      import :: c_long, c_int, c_double
      integer(kind = c_long) :: nnodes_a, nelements_a, nnodes_b, nelements_b
      integer(kind = c_int) :: dim_a, loc_a, dim_b, loc_b
      real(kind = c_double), dimension(*) :: positions_a, positions_b
      integer(kind = c_long), dimension(*) :: enlist_a, enlist_b
    end subroutine c_tree_set_input
    subroutine c_tree_query_output(nindices) &
        & bind(c, name = "libsupermesh_tree_intersection_finder_query_output")
      import :: c_long
      integer(kind = c_long) :: nindices
    end subroutine c_tree_query_output
    subroutine c_tree_get_output(nelements, nindices, indices, ind_ptr) &
        & bind(c, name = "libsupermesh_tree_intersection_finder_get_output")
      import :: c_long
      integer(kind = c_long) :: nelements, nindices
      integer(kind = c_long), dimension(*) :: indices, ind_ptr
    end subroutine c_tree_get_output
    subroutine c_rtree_set_input(nnodes_a, dim_a, nelements_a, loc_a, nnodes_b, dim_b, &
        & nelements_b, loc_b, positions_a, enlist_a, positions_b, enlist_b) &
        & bind(c, name = "libsupermesh_rtree_intersection_finder_set_input")
      import :: c_long, c_int, c_double
      integer(kind = c_long) :: nnodes_a, nelements_a, nnodes_b, nelements_b
      integer(kind = c_int) :: dim_a, loc_a, dim_b, loc_b
      real(kind = c_double), dimension(*) :: positions_a, positions_b
      integer(kind = c_long), dimension(*) :: enlist_a, enlist_b
    end subroutine c_rtree_set_input
    subroutine c_rtree_query_output(nindices) &
        & bind(c, name = "libsupermesh_rtree_intersection_finder_query_output")
      import :: c_long
      integer(kind = c_long) :: nindices
    end subroutine c_rtree_query_output
    subroutine c_rtree_get_output(nelements, nindices, indices, ind_ptr) &
        & bind(c, name = "libsupermesh_rtree_intersection_finder_get_output")
      import :: c_long
      integer(kind = c_long) :: nelements, nindices
      integer(kind = c_long), dimension(*) :: indices, ind_ptr
    end subroutine c_rtree_get_output
    subroutine c_front_set_input(nnodes_a, dim_a, nelements_a, loc_a, nnodes_b, dim_b, &
        & nelements_b, loc_b, positions_a, enlist_a, positions_b, enlist_b) &
        & bind(c, name = "libsupermesh_advancing_front_intersection_finder_set_input")
      import :: c_long, c_int, c_double
      integer(kind = c_long) :: nnodes_a, nelements_a, nnodes_b, nelements_b
      integer(kind = c_int) :: dim_a, loc_a, dim_b, loc_b
      real(kind = c_double), dimension(*) :: positions_a, positions_b
      integer(kind = c_long), dimension(*) :: enlist_a, enlist_b
    end subroutine c_front_set_input
    subroutine c_front_query_output(nindices) &
        & bind(c, name = "libsupermesh_advancing_front_intersection_finder_query_output")
      import :: c_long
      integer(kind = c_long) :: nindices
    end subroutine c_front_query_output
    subroutine c_front_get_output(nelements, nindices, indices, ind_ptr) &
        & bind(c, name = "libsupermesh_advancing_front_intersection_finder_get_output")
      import :: c_long
      integer(kind = c_long) :: nelements, nindices
      integer(kind = c_long), dimension(*) :: indices, ind_ptr
    end subroutine c_front_get_output
    subroutine c_intersect_tris_real(tri_a, tri_b, tris_c, n_tris_c) &
        & bind(c, name = "libsupermesh_intersect_tris_real")
      import :: c_double, c_int
      real(kind = c_double), dimension(*) :: tri_a, tri_b, tris_c
      integer(kind = c_int) :: n_tris_c
    end subroutine c_intersect_tris_real
    subroutine c_intersect_tets_real(tet_a, tet_b, tets_c, n_tets_c) &
        & bind(c, name = "libsupermesh_intersect_tets_real")
      import :: c_double, c_int
      real(kind = c_double), dimension(*) :: tet_a, tet_b, tets_c
      integer(kind = c_int) :: n_tets_c
    end subroutine c_intersect_tets_real
    subroutine c_triangle_area(tri, area) bind(c, name = "libsupermesh_triangle_area")
      import :: c_double
      real(kind = c_double), dimension(*) :: tri
      real(kind = c_double) :: area
    end subroutine c_triangle_area
    subroutine c_tetrahedron_volume(tet, volume) bind(c, name = "libsupermesh_tetrahedron_volume")
      import :: c_double
      real(kind = c_double), dimension(*) :: tet
      real(kind = c_double) :: volume
    end subroutine c_tetrahedron_volume
    subroutine c_sort_set_input(nnodes_a, dim_a, nelements_a, loc_a, nnodes_b, dim_b, &
        & nelements_b, loc_b, positions_a, enlist_a, positions_b, enlist_b) &
        & bind(c, name = "libsupermesh_sort_intersection_finder_set_input")
      import :: c_long, c_int, c_double
      integer(kind = c_long) :: nnodes_a, nelements_a, nnodes_b, nelements_b
      integer(kind = c_int) :: dim_a, loc_a, dim_b, loc_b
      real(kind = c_double), dimension(*) :: positions_a, positions_b
      integer(kind = c_long), dimension(*) :: enlist_a, enlist_b
    end subroutine c_sort_set_input
    subroutine c_sort_query_output(nindices) &
        & bind(c, name = "libsupermesh_sort_intersection_finder_query_output")
      import :: c_long
      integer(kind = c_long) :: nindices
    end subroutine c_sort_query_output
    subroutine c_sort_get_output(nelements, nindices, indices, ind_ptr) &
        & bind(c, name = "libsupermesh_sort_intersection_finder_get_output")
      import :: c_long
      integer(kind = c_long) :: nelements, nindices
      integer(kind = c_long), dimension(*) :: indices, ind_ptr
    end subroutine c_sort_get_output
    subroutine c_intersect_tri_quad(tri_a, quad_b, tris_c, n_tris_c) &
        & bind(c, name = "libsupermesh_intersect_tri_quad")
      import :: c_double, c_int
      real(kind = c_double), dimension(*) :: tri_a, quad_b, tris_c
      integer(kind = c_int) :: n_tris_c
    end subroutine c_intersect_tri_quad
    subroutine c_intersect_tet_hex(tet_a, hex_b, tets_c, n_tets_c) &
        & bind(c, name = "libsupermesh_intersect_tet_hex")
      import :: c_double, c_int
      real(kind = c_double), dimension(*) :: tet_a, hex_b, tets_c
      integer(kind = c_int) :: n_tets_c
    end subroutine c_intersect_tet_hex
    subroutine c_intersect_quads(quad_a, quad_b, tris_c, n_tris_c) &
        & bind(c, name = "libsupermesh_intersect_quads")
      import :: c_double, c_int
      real(kind = c_double), dimension(*) :: quad_a, quad_b, tris_c
      integer(kind = c_int) :: n_tris_c
    end subroutine c_intersect_quads
    subroutine c_intersect_hexes(hex_a, hex_b, tets_c, n_tets_c) &
        & bind(c, name = "libsupermesh_intersect_hexes")
      import :: c_double, c_int
      real(kind = c_double), dimension(*) :: hex_a, hex_b, tets_c
      integer(kind = c_int) :: n_tets_c
    end subroutine c_intersect_hexes
    subroutine c_intersect_intervals(interval_a, interval_b, intervals_c, n_intervals_c) &
        & bind(c, name = "libsupermesh_intersect_intervals")
      import :: c_double, c_int
      real(kind = c_double), dimension(*) :: interval_a, interval_b, intervals_c
      integer(kind = c_int) :: n_intervals_c
    end subroutine c_intersect_intervals
    subroutine c_interval_size(interval, size) bind(c, name = "libsupermesh_interval_size")
      import :: c_double
      real(kind = c_double), dimension(*) :: interval
      real(kind = c_double) :: size
    end subroutine c_interval_size
    function c_context_create(device, ctx) bind(c, name = "libsupermesh_gpu_context_create")
      import :: c_int, c_ptr
      integer(kind = c_int), value :: device
      type(c_ptr) :: ctx
      integer(kind = c_int) :: c_context_create
    end function c_context_create
    function c_gpu_malloc(device, bytes, ptr) bind(c, name = "libsupermesh_gpu_malloc")
      import :: c_int, c_size_t, c_ptr
      integer(kind = c_int), value :: device
      integer(kind = c_size_t), value :: bytes
      type(c_ptr) :: ptr
      integer(kind = c_int) :: c_gpu_malloc
    end function c_gpu_malloc
    function c_gpu_free(ptr) bind(c, name = "libsupermesh_gpu_free")
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
      integer(kind = c_int) :: c_gpu_free
    end function c_gpu_free
    function c_gpu_memcpy(dst, src, bytes, kind) bind(c, name = "libsupermesh_gpu_memcpy")
      import :: c_int, c_size_t, c_ptr
      type(c_ptr), value :: dst, src
      integer(kind = c_size_t), value :: bytes
      integer(kind = c_int), value :: kind
      integer(kind = c_int) :: c_gpu_memcpy
    end function c_gpu_memcpy
    function c_supermesh_callback(ctx, a, b, callback, user) &
        & bind(c, name = "libsupermesh_gpu_supermesh_callback")
      import :: c_int, c_ptr, c_funptr, gpu_mesh
      type(c_ptr), value :: ctx
      type(gpu_mesh) :: a, b
      type(c_funptr), value :: callback
      type(c_ptr), value :: user
      integer(kind = c_int) :: c_supermesh_callback
    end function c_supermesh_callback
    function c_supermesh_callback_mixed(ctx, a, loc_a, b, loc_b, callback, user) &
        & bind(c, name = "libsupermesh_gpu_supermesh_callback_mixed")
      import :: c_int, c_ptr, c_funptr, gpu_mesh
      type(c_ptr), value :: ctx
      type(gpu_mesh) :: a, b
      integer(kind = c_int), value :: loc_a, loc_b
      type(c_funptr), value :: callback
      type(c_ptr), value :: user
      integer(kind = c_int) :: c_supermesh_callback_mixed
    end function c_supermesh_callback_mixed
    function c_last_error() bind(c, name = "libsupermesh_gpu_last_error")
      import :: c_ptr
      type(c_ptr) :: c_last_error
    end function c_last_error
    function c_last_error_code() bind(c, name = "libsupermesh_gpu_last_error_code")
      import :: c_int
      integer(kind = c_int) :: c_last_error_code
    end function c_last_error_code
    function c_tree_create(finder, dim, nnodes, positions, nelements, loc, enlist, tree) &
        & bind(c, name = "libsupermesh_gpu_tree_create")
      import :: c_int, c_long, c_double, c_ptr
      integer(kind = c_int), value :: finder, dim, loc
      integer(kind = c_long), value :: nnodes, nelements
      real(kind = c_double), dimension(*) :: positions
      integer(kind = c_long), dimension(*) :: enlist
      type(c_ptr) :: tree
      integer(kind = c_int) :: c_tree_create
    end function c_tree_create
    function c_tree_query(tree, loc, element, neles) bind(c, name = "libsupermesh_gpu_tree_query")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: tree
      integer(kind = c_int), value :: loc
      real(kind = c_double), dimension(*) :: element
      integer(kind = c_long) :: neles
      integer(kind = c_int) :: c_tree_query
    end function c_tree_query
    function c_tree_result(tree, neles, eles) bind(c, name = "libsupermesh_gpu_tree_result")
      import :: c_int, c_long, c_ptr
      type(c_ptr), value :: tree
      integer(kind = c_long), value :: neles
      integer(kind = c_long), dimension(*) :: eles
      integer(kind = c_int) :: c_tree_result
    end function c_tree_result
    subroutine c_tree_destroy(tree) bind(c, name = "libsupermesh_gpu_tree_destroy")
      import :: c_ptr
      type(c_ptr), value :: tree
    end subroutine c_tree_destroy
  end interface

contains

  ! Message of the last failed call on this thread ("" after a success).
  function libsupermesh_gpu_error() result(msg)
    character(len = :), allocatable :: msg
    character(kind = c_char), dimension(:), pointer :: chars
    type(c_ptr) :: p
    integer :: n
    p = c_last_error()
    if(.not. c_associated(p)) then
      msg = ""
      return
    end if
    call c_f_pointer(p, chars, [4096])
    n = 0
    do while(n < 4096)
      if(chars(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len = n) :: msg)
    msg = transfer(chars(:n), msg)
  end function libsupermesh_gpu_error

  ! CSR candidate pairs, strict AABB overlap (Intersection_Finder.F90:785-809).
  subroutine tree_intersection_finder_csr_sparsity(positions_a, enlist_a, positions_b, enlist_b, &
      & map_ab_indices, map_ab_indptr)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    integer, dimension(:), allocatable, intent(out) :: map_ab_indices
    integer, dimension(:), intent(out) :: map_ab_indptr
    call finder_csr(finder_tree, positions_a, enlist_a, positions_b, enlist_b, map_ab_indices, &
      & map_ab_indptr)
  end subroutine tree_intersection_finder_csr_sparsity

  ! CSR candidate pairs, closed AABB overlap (Intersection_Finder.F90:621-643).
  subroutine rtree_intersection_finder_csr_sparsity(positions_a, enlist_a, positions_b, enlist_b, &
      & map_ab_indices, map_ab_indptr)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    integer, dimension(:), allocatable, intent(out) :: map_ab_indices
    integer, dimension(:), intent(out) :: map_ab_indptr
    call finder_csr(finder_rtree, positions_a, enlist_a, positions_b, enlist_b, map_ab_indices, &
      & map_ab_indptr)
  end subroutine rtree_intersection_finder_csr_sparsity

  ! CSR candidate pairs of the advancing front finder
  ! (Intersection_Finder.F90:270-476): closed overlap, rows in the front's
  ! visiting order.
  subroutine advancing_front_intersection_finder_csr_sparsity(positions_a, enlist_a, positions_b, &
      & enlist_b, map_ab_indices, map_ab_indptr)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    integer, dimension(:), allocatable, intent(out) :: map_ab_indices
    integer, dimension(:), intent(out) :: map_ab_indptr
    call finder_csr(finder_front, positions_a, enlist_a, positions_b, enlist_b, map_ab_indices, &
      & map_ab_indptr)
  end subroutine advancing_front_intersection_finder_csr_sparsity

  ! The default finder by dimension (Intersection_Finder.F90:234-262): the
  ! sort finder in 1D, the R-tree finder in 2D and 3D.
  subroutine intersection_finder_csr_sparsity(positions_a, enlist_a, positions_b, enlist_b, &
      & map_ab_indices, map_ab_indptr)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    integer, dimension(:), allocatable, intent(out) :: map_ab_indices
    integer, dimension(:), intent(out) :: map_ab_indptr
    integer :: kind
    kind = finder_rtree
    if(size(positions_a, 1) == 1) kind = finder_sort
    call finder_csr(kind, positions_a, enlist_a, positions_b, enlist_b, map_ab_indices, &
      & map_ab_indptr)
  end subroutine intersection_finder_csr_sparsity

  subroutine finder_csr(kind, positions_a, enlist_a, positions_b, enlist_b, map_ab_indices, &
      & map_ab_indptr)
    integer, intent(in) :: kind
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    integer, dimension(:), allocatable, intent(out) :: map_ab_indices
    integer, dimension(:), intent(out) :: map_ab_indptr
    integer(kind = c_long) :: nn_a, ne_a, nn_b, ne_b, nind
    integer(kind = c_int) :: dim_a, loc_a, dim_b, loc_b
    integer(kind = c_long), dimension(:, :), allocatable :: ea, eb
    integer(kind = c_long), dimension(:), allocatable :: ind, ptr
    real(kind = c_double), dimension(:, :), allocatable :: pa, pb
    dim_a = size(positions_a, 1)
    nn_a = size(positions_a, 2)
    loc_a = size(enlist_a, 1)
    ne_a = size(enlist_a, 2)
    dim_b = size(positions_b, 1)
    nn_b = size(positions_b, 2)
    loc_b = size(enlist_b, 1)
    ne_b = size(enlist_b, 2)
    allocate(pa(dim_a, nn_a), pb(dim_b, nn_b), ea(loc_a, ne_a), eb(loc_b, ne_b))
    pa = positions_a
    pb = positions_b
    ea = enlist_a - 1
    eb = enlist_b - 1
    select case(kind)
      case(finder_tree)
        call c_tree_set_input(nn_a, dim_a, ne_a, loc_a, nn_b, dim_b, ne_b, loc_b, pa, ea, pb, eb)
      case(finder_rtree)
        call c_rtree_set_input(nn_a, dim_a, ne_a, loc_a, nn_b, dim_b, ne_b, loc_b, pa, ea, pb, eb)
      case(finder_sort)
        call c_sort_set_input(nn_a, dim_a, ne_a, loc_a, nn_b, dim_b, ne_b, loc_b, pa, ea, pb, eb)
      case default
        call c_front_set_input(nn_a, dim_a, ne_a, loc_a, nn_b, dim_b, ne_b, loc_b, pa, ea, pb, eb)
    end select
    nind = 0
    if(c_last_error_code() == 0) then
      select case(kind)
        case(finder_tree)
          call c_tree_query_output(nind)
        case(finder_rtree)
          call c_rtree_query_output(nind)
        case(finder_sort)
          call c_sort_query_output(nind)
        case default
          call c_front_query_output(nind)
      end select
    end if
    allocate(ind(max(nind, 1_c_long)), ptr(ne_a + 1))
    ptr = 0
    if(c_last_error_code() == 0) then
      select case(kind)
        case(finder_tree)
          call c_tree_get_output(ne_a, nind, ind, ptr)
        case(finder_rtree)
          call c_rtree_get_output(ne_a, nind, ind, ptr)
        case(finder_sort)
          call c_sort_get_output(ne_a, nind, ind, ptr)
        case default
          call c_front_get_output(ne_a, nind, ind, ptr)
      end select
    end if
    allocate(map_ab_indices(nind))
    map_ab_indices = int(ind(:nind)) + 1
    map_ab_indptr = int(ptr) + 1
  end subroutine finder_csr

  ! CSR candidate pairs of the 1D sort finder (Intersection_Finder.F90:924-948).
  subroutine sort_intersection_finder_csr_sparsity(positions_a, enlist_a, positions_b, enlist_b, &
      & map_ab_indices, map_ab_indptr)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    integer, dimension(:), allocatable, intent(out) :: map_ab_indices
    integer, dimension(:), intent(out) :: map_ab_indptr
    call finder_csr(finder_sort, positions_a, enlist_a, positions_b, enlist_b, map_ab_indices, &
      & map_ab_indptr)
  end subroutine sort_intersection_finder_csr_sparsity

  ! intersect_intervals (Interval_Intersection.F90:45-70).
  subroutine intersect_intervals(interval_a, interval_b, intervals_c, n_intervals_c)
    real(kind = real_kind), dimension(2), intent(in) :: interval_a, interval_b
    real(kind = real_kind), dimension(2), intent(out) :: intervals_c
    integer, intent(out) :: n_intervals_c
    real(kind = c_double), dimension(2) :: a, b
    integer(kind = c_int) :: n
    a = interval_a
    b = interval_b
    call c_intersect_intervals(a, b, intervals_c, n)
    n_intervals_c = n
  end subroutine intersect_intervals

  function interval_size(interval) result(size)
    real(kind = real_kind), dimension(2), intent(in) :: interval
    real(kind = real_kind) :: size
    real(kind = c_double), dimension(2) :: v
    v = interval
    call c_interval_size(v, size)
  end function interval_size

  ! intersect_tri_quad (Supermesh.F90:211-230): tris_c(2, 3, 46).
  subroutine intersect_tri_quad(tri_a, quad_b, tris_c, n_tris_c)
    real(kind = real_kind), dimension(2, 3), intent(in) :: tri_a
    real(kind = real_kind), dimension(2, 4), intent(in) :: quad_b
    real(kind = real_kind), dimension(:, :, :), intent(inout) :: tris_c
    integer, intent(out) :: n_tris_c
    real(kind = c_double), dimension(2, 3) :: a
    real(kind = c_double), dimension(2, 4) :: b
    real(kind = c_double), dimension(2, 3, 46) :: c
    integer(kind = c_int) :: n
    a = tri_a
    b = quad_b
    call c_intersect_tri_quad(a, b, c, n)
    n_tris_c = n
    if(n > 0) tris_c(:, :, :n) = c(:, :, :n)
  end subroutine intersect_tri_quad

  ! intersect_tet_hex (Supermesh.F90:251-301): tets_c(3, 4, 729).
  subroutine intersect_tet_hex(tet_a, hex_b, tets_c, n_tets_c)
    real(kind = real_kind), dimension(3, 4), intent(in) :: tet_a
    real(kind = real_kind), dimension(3, 8), intent(in) :: hex_b
    real(kind = real_kind), dimension(:, :, :), intent(inout) :: tets_c
    integer, intent(out) :: n_tets_c
    real(kind = c_double), dimension(3, 4) :: a
    real(kind = c_double), dimension(3, 8) :: b
    real(kind = c_double), dimension(:, :, :), allocatable :: c
    integer(kind = c_int) :: n
    allocate(c(3, 4, 729))
    a = tet_a
    b = hex_b
    call c_intersect_tet_hex(a, b, c, n)
    n_tets_c = n
    if(n > 0) tets_c(:, :, :n) = c(:, :, :n)
  end subroutine intersect_tet_hex

  ! Callback driver: the single-GPU form of the self-self loop of
  ! parallel_supermesh (Parallel_Supermesh.F90:436-453) without the MPI
  ! exchange.  intersection_calculation is called for every closed-overlap
  ! candidate pair, B element by B element, with local = .true.  Elements may
  ! be simplices, quadrilaterals or hexahedra (loc = size(enlist, 1)); each
  ! pair is clipped as intersect_elements does.  On failure (bad node ids,
  ! element types, device errors) no callback is made past the failure and
  ! libsupermesh_gpu_error() holds the message.
  subroutine supermesh(positions_a, enlist_a, positions_b, enlist_b, intersection_calculation)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    procedure(intersection_calculation_interface) :: intersection_calculation
    type(gpu_mesh) :: ma, mb
    type(c_ptr) :: bufs(4)
    integer(kind = c_int) :: rc, rc_free
    integer :: i
    bufs = c_null_ptr
    if(.not. c_associated(gpu_context)) then
      rc = c_context_create(0_c_int, gpu_context)
      if(rc /= 0) return
    end if
    call upload(positions_a, enlist_a, ma, bufs(1), bufs(2), rc)
    if(rc == 0) call upload(positions_b, enlist_b, mb, bufs(3), bufs(4), rc)
    if(rc == 0) then
      user_calculation => intersection_calculation
      callback_dim = size(positions_a, 1)
      callback_loc_a = size(enlist_a, 1)
      callback_loc_b = size(enlist_b, 1)
      rc = c_supermesh_callback_mixed(gpu_context, ma, int(callback_loc_a, c_int), mb, &
        & int(callback_loc_b, c_int), c_funloc(callback_trampoline), c_null_ptr)
      user_calculation => null()
    end if
    ! a successful free leaves the message of a failed call above readable
    do i = 1, 4
      if(c_associated(bufs(i))) rc_free = c_gpu_free(bufs(i))
    end do
  end subroutine supermesh

  subroutine upload(positions, enlist, mesh, dpos, denl, rc)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions
    integer, dimension(:, :), intent(in) :: enlist
    type(gpu_mesh), intent(out) :: mesh
    type(c_ptr), intent(inout) :: dpos, denl
    integer(kind = c_int), intent(out) :: rc
    real(kind = c_double), dimension(:, :), allocatable, target :: p
    integer(kind = c_int), dimension(:, :), allocatable, target :: e
    integer(kind = c_size_t) :: np, ne
    allocate(p(size(positions, 1), size(positions, 2)), e(size(enlist, 1), size(enlist, 2)))
    p = positions
    e = enlist - 1
    np = 8_c_size_t * max(size(p), 1)
    ne = 4_c_size_t * max(size(e), 1)
    mesh%dim = size(positions, 1)
    mesh%nnodes = size(positions, 2)
    mesh%nelements = size(enlist, 2)
    rc = c_gpu_malloc(0_c_int, np, dpos)
    if(rc == 0) rc = c_gpu_malloc(0_c_int, ne, denl)
    if(rc == 0 .and. size(p) > 0) rc = c_gpu_memcpy(dpos, c_loc(p), 8_c_size_t * size(p), 1_c_int)
    if(rc == 0 .and. size(e) > 0) rc = c_gpu_memcpy(denl, c_loc(e), 4_c_size_t * size(e), 1_c_int)
    mesh%positions = dpos
    mesh%enlist = denl
  end subroutine upload

  subroutine callback_trampoline(element_a, element_b, elements_c, n_elements_c, nodes_b, &
      & ele_a, ele_b, user) bind(c)
    type(c_ptr), value :: element_a, element_b, elements_c, nodes_b, user
    integer(kind = c_int), value :: n_elements_c
    integer(kind = c_long), value :: ele_a, ele_b
    real(kind = c_double), dimension(:, :), pointer :: ea, eb
    real(kind = c_double), dimension(:, :, :), pointer :: ec
    real(kind = c_double), dimension(callback_dim, callback_dim + 1, 0), target :: empty
    integer(kind = c_long), dimension(:), pointer :: nb
    integer :: d
    d = callback_dim
    call c_f_pointer(element_a, ea, [d, callback_loc_a])
    call c_f_pointer(element_b, eb, [d, callback_loc_b])
    call c_f_pointer(nodes_b, nb, [callback_loc_b])
    if(n_elements_c > 0) then
      call c_f_pointer(elements_c, ec, [d, d + 1, int(n_elements_c)])
    else
      ec => empty
    end if
    call user_calculation(ea, eb, ec, int(nb) + 1, int(ele_a) + 1, int(ele_b) + 1, .true.)
  end subroutine callback_trampoline

  ! intersect_tris_real (Triangle_Intersection.F90:125-142).
  subroutine intersect_tris_real(tri_a, tri_b, tris_c, n_tris_c)
    real(kind = real_kind), dimension(2, 3), intent(in) :: tri_a, tri_b
    real(kind = real_kind), dimension(2, 3, tri_buf_size), intent(out) :: tris_c
    integer, intent(out) :: n_tris_c
    real(kind = c_double), dimension(2, 3) :: a, b
    integer(kind = c_int) :: n
    a = tri_a
    b = tri_b
    call c_intersect_tris_real(a, b, tris_c, n)
    n_tris_c = n
  end subroutine intersect_tris_real

  ! intersect_tets_real (Tetrahedron_Intersection.F90:107-125).
  subroutine intersect_tets_real(tet_a, tet_b, tets_c, n_tets_c)
    real(kind = real_kind), dimension(3, 4), intent(in) :: tet_a, tet_b
    real(kind = real_kind), dimension(3, 4, tet_buf_size), intent(out) :: tets_c
    integer, intent(out) :: n_tets_c
    real(kind = c_double), dimension(3, 4) :: a, b
    integer(kind = c_int) :: n
    a = tet_a
    b = tet_b
    call c_intersect_tets_real(a, b, tets_c, n)
    n_tets_c = n
  end subroutine intersect_tets_real

  function triangle_area(tri) result(area)
    real(kind = real_kind), dimension(2, 3), intent(in) :: tri
    real(kind = real_kind) :: area
    real(kind = c_double), dimension(2, 3) :: t
    t = tri
    call c_triangle_area(t, area)
  end function triangle_area

  function tetrahedron_volume(tet) result(volume)
    real(kind = real_kind), dimension(3, 4), intent(in) :: tet
    real(kind = real_kind) :: volume
    real(kind = c_double), dimension(3, 4) :: t
    t = tet
    call c_tetrahedron_volume(t, volume)
  end function tetrahedron_volume


  ! ---------------------------------------------------------------------
  ! Device trees for one-element queries (Intersection_Finder.F90:481-760).
  ! Results are ascending B element ids (1-based).
  ! ---------------------------------------------------------------------

  subroutine tree_create(finder, positions, enlist, handle)
    integer, intent(in) :: finder
    real(kind = real_kind), dimension(:, :), intent(in) :: positions
    integer, dimension(:, :), intent(in) :: enlist
    type(c_ptr), intent(out) :: handle
    integer(kind = c_long), dimension(:, :), allocatable :: e
    real(kind = c_double), dimension(:, :), allocatable :: p
    integer(kind = c_int) :: rc
    allocate(p(size(positions, 1), size(positions, 2)), e(size(enlist, 1), size(enlist, 2)))
    p = positions
    e = enlist - 1
    handle = c_null_ptr
    rc = c_tree_create(int(finder, c_int), int(size(positions, 1), c_int), int(size(positions, 2), c_long), &
      & p, int(size(enlist, 2), c_long), int(size(enlist, 1), c_int), e, handle)
  end subroutine tree_create

  subroutine tree_query(handle, element_a, eles_b)
    type(c_ptr), intent(in) :: handle
    real(kind = real_kind), dimension(:, :), intent(in) :: element_a
    integer, dimension(:), intent(out), allocatable :: eles_b
    real(kind = c_double), dimension(size(element_a, 1), size(element_a, 2)) :: el
    integer(kind = c_long), dimension(:), allocatable :: r
    integer(kind = c_long) :: n
    integer(kind = c_int) :: rc
    el = element_a
    n = 0
    if(c_associated(handle)) rc = c_tree_query(handle, int(size(element_a, 2), c_int), el, n)
    allocate(r(max(n, 1_c_long)))
    if(n > 0) rc = c_tree_result(handle, n, r)
    allocate(eles_b(n))
    eles_b = int(r(:n)) + 1
  end subroutine tree_query

  subroutine allocate_tree(tree, positions, enlist)
    type(tree_type), intent(out) :: tree
    real(kind = real_kind), dimension(:, :), intent(in) :: positions
    integer, dimension(:, :), intent(in) :: enlist
    tree%dim = size(positions, 1)
    call tree_create(0, positions, enlist, tree%handle)
  end subroutine allocate_tree

  subroutine allocate_rtree(rtree, positions, enlist)
    type(rtree_type), intent(out) :: rtree
    real(kind = real_kind), dimension(:, :), intent(in) :: positions
    integer, dimension(:, :), intent(in) :: enlist
    call tree_create(1, positions, enlist, rtree%handle)
  end subroutine allocate_rtree

  subroutine query_tree_allocatable(tree, element_a, eles_b)
    type(tree_type), intent(inout) :: tree
    real(kind = real_kind), dimension(:, :), intent(in) :: element_a
    integer, dimension(:), allocatable, intent(out) :: eles_b
    call tree_query(tree%handle, element_a, eles_b)
  end subroutine query_tree_allocatable

  subroutine query_tree_pointer(tree, element_a, eles_b)
    type(tree_type), intent(inout) :: tree
    real(kind = real_kind), dimension(:, :), intent(in) :: element_a
    integer, dimension(:), pointer, intent(out) :: eles_b
    integer, dimension(:), allocatable :: r
    call tree_query(tree%handle, element_a, r)
    allocate(eles_b(size(r)))
    eles_b = r
  end subroutine query_tree_pointer

  subroutine query_rtree_allocatable(rtree, element_a, eles_b)
    type(rtree_type), intent(inout) :: rtree
    real(kind = real_kind), dimension(:, :), intent(in) :: element_a
    integer, dimension(:), allocatable, intent(out) :: eles_b
    call tree_query(rtree%handle, element_a, eles_b)
  end subroutine query_rtree_allocatable

  subroutine query_rtree_pointer(rtree, element_a, eles_b)
    type(rtree_type), intent(inout) :: rtree
    real(kind = real_kind), dimension(:, :), intent(in) :: element_a
    integer, dimension(:), pointer, intent(out) :: eles_b
    integer, dimension(:), allocatable :: r
    call tree_query(rtree%handle, element_a, r)
    allocate(eles_b(size(r)))
    eles_b = r
  end subroutine query_rtree_pointer

  subroutine deallocate_tree(tree)
    type(tree_type), intent(inout) :: tree
    if(c_associated(tree%handle)) call c_tree_destroy(tree%handle)
    tree%handle = c_null_ptr
  end subroutine deallocate_tree

  subroutine deallocate_rtree(rtree)
    type(rtree_type), intent(inout) :: rtree
    if(c_associated(rtree%handle)) call c_tree_destroy(rtree%handle)
    rtree%handle = c_null_ptr
  end subroutine deallocate_rtree

  subroutine tree_intersection_finder_set_input(positions, enlist)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions
    integer, dimension(:, :), intent(in) :: enlist
    call tree_intersection_finder_reset()
    call allocate(state_tree, positions, enlist)
  end subroutine tree_intersection_finder_set_input

  subroutine tree_intersection_finder_find(element_a)
    real(kind = real_kind), dimension(:, :), intent(in) :: element_a
    if(allocated(state_tree_eles)) deallocate(state_tree_eles)
    call query(state_tree, element_a, state_tree_eles)
  end subroutine tree_intersection_finder_find

  subroutine tree_intersection_finder_query_output(neles_b)
    integer, intent(out) :: neles_b
    neles_b = 0
    if(allocated(state_tree_eles)) neles_b = size(state_tree_eles)
  end subroutine tree_intersection_finder_query_output

  subroutine tree_intersection_finder_get_output(ele_b, i)
    integer, intent(out) :: ele_b
    integer, intent(in) :: i
    ele_b = state_tree_eles(i)
  end subroutine tree_intersection_finder_get_output

  subroutine tree_intersection_finder_reset()
    if(allocated(state_tree_eles)) deallocate(state_tree_eles)
    call deallocate(state_tree)
  end subroutine tree_intersection_finder_reset

  subroutine rtree_intersection_finder_set_input(positions, enlist)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions
    integer, dimension(:, :), intent(in) :: enlist
    call rtree_intersection_finder_reset()
    call allocate(state_rtree, positions, enlist)
  end subroutine rtree_intersection_finder_set_input

  subroutine rtree_intersection_finder_find(element_a)
    real(kind = real_kind), dimension(:, :), intent(in) :: element_a
    if(allocated(state_rtree_eles)) deallocate(state_rtree_eles)
    call query(state_rtree, element_a, state_rtree_eles)
  end subroutine rtree_intersection_finder_find

  subroutine rtree_intersection_finder_query_output(neles_b)
    integer, intent(out) :: neles_b
    neles_b = 0
    if(allocated(state_rtree_eles)) neles_b = size(state_rtree_eles)
  end subroutine rtree_intersection_finder_query_output

  subroutine rtree_intersection_finder_get_output(ele_b, i)
    integer, intent(out) :: ele_b
    integer, intent(in) :: i
    ele_b = state_rtree_eles(i)
  end subroutine rtree_intersection_finder_get_output

  subroutine rtree_intersection_finder_reset()
    if(allocated(state_rtree_eles)) deallocate(state_rtree_eles)
    call deallocate(state_rtree)
  end subroutine rtree_intersection_finder_reset

  ! ---------------------------------------------------------------------
  ! Ragged (intersections) forms of the finders: the CSR result split per
  ! A element (Intersections.F90:48-71 run backwards).
  ! ---------------------------------------------------------------------

  subroutine csr_to_intersections(indices, indptr, map_ab)
    integer, dimension(:), intent(in) :: indices, indptr
    type(intersections), dimension(:), intent(out) :: map_ab
    integer :: a, n
    do a = 1, size(map_ab)
      n = indptr(a + 1) - indptr(a)
      allocate(map_ab(a)%v(n))
      map_ab(a)%v = indices(indptr(a):indptr(a + 1) - 1)
      map_ab(a)%n = n
    end do
  end subroutine csr_to_intersections

  subroutine deallocate_intersections(ints)
    type(intersections), intent(inout) :: ints
    if(associated(ints%v)) deallocate(ints%v)
    ints%n = 0
  end subroutine deallocate_intersections

  subroutine deallocate_intersections_rank_1(ints)
    type(intersections), dimension(:), intent(inout) :: ints
    integer :: i
    do i = 1, size(ints)
      call deallocate_intersections(ints(i))
    end do
  end subroutine deallocate_intersections_rank_1

  subroutine ragged(kind, positions_a, enlist_a, positions_b, enlist_b, map_ab)
    integer, intent(in) :: kind
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    type(intersections), dimension(:), intent(out) :: map_ab
    integer, dimension(:), allocatable :: indices
    integer, dimension(size(enlist_a, 2) + 1) :: indptr
    call finder_csr(kind, positions_a, enlist_a, positions_b, enlist_b, indices, indptr)
    call csr_to_intersections(indices, indptr, map_ab)
  end subroutine ragged

  subroutine tree_intersection_finder_intersections(positions_a, enlist_a, positions_b, enlist_b, map_ab)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    type(intersections), dimension(:), intent(out) :: map_ab
    call ragged(finder_tree, positions_a, enlist_a, positions_b, enlist_b, map_ab)
  end subroutine tree_intersection_finder_intersections

  subroutine rtree_intersection_finder_intersections(positions_a, enlist_a, positions_b, enlist_b, map_ab)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    type(intersections), dimension(:), intent(out) :: map_ab
    call ragged(finder_rtree, positions_a, enlist_a, positions_b, enlist_b, map_ab)
  end subroutine rtree_intersection_finder_intersections

  subroutine sort_intersection_finder_intersections(positions_a, enlist_a, positions_b, enlist_b, map_ab)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    type(intersections), dimension(:), intent(out) :: map_ab
    call ragged(finder_sort, positions_a, enlist_a, positions_b, enlist_b, map_ab)
  end subroutine sort_intersection_finder_intersections

  subroutine advancing_front_intersection_finder_intersections(positions_a, enlist_a, positions_b, &
      & enlist_b, map_ab)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    type(intersections), dimension(:), intent(out) :: map_ab
    call ragged(finder_front, positions_a, enlist_a, positions_b, enlist_b, map_ab)
  end subroutine advancing_front_intersection_finder_intersections

  subroutine intersection_finder_intersections(positions_a, enlist_a, positions_b, enlist_b, map_ab)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    type(intersections), dimension(:), intent(out) :: map_ab
    integer :: kind
    kind = finder_rtree
    if(size(positions_a, 1) == 1) kind = finder_sort
    call ragged(kind, positions_a, enlist_a, positions_b, enlist_b, map_ab)
  end subroutine intersection_finder_intersections

  ! Brute force (Intersection_Finder.F90:990-1050): the strict-overlap set,
  ! rows ascending -- the tree finder's result exactly.
  subroutine brute_force_intersection_finder_csr_sparsity(positions_a, enlist_a, positions_b, enlist_b, &
      & map_ab_indices, map_ab_indptr)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    integer, dimension(:), allocatable, intent(out) :: map_ab_indices
    integer, dimension(:), intent(out) :: map_ab_indptr
    call finder_csr(finder_tree, positions_a, enlist_a, positions_b, enlist_b, map_ab_indices, &
      & map_ab_indptr)
  end subroutine brute_force_intersection_finder_csr_sparsity

  subroutine brute_force_intersection_finder_intersections(positions_a, enlist_a, positions_b, enlist_b, &
      & map_ab)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    type(intersections), dimension(:), intent(out) :: map_ab
    call ragged(finder_tree, positions_a, enlist_a, positions_b, enlist_b, map_ab)
  end subroutine brute_force_intersection_finder_intersections


  ! ---------------------------------------------------------------------
  ! Element dispatch (Supermesh.F90:89-209, 231-248, 303-351) and the buffer
  ! sizes of the per-pair calls (Triangle_Intersection.F90:106-123,
  ! Tetrahedron_Intersection.F90:181-188).
  ! ---------------------------------------------------------------------

  pure elemental function max_n_tris_c_tri(n_lines_b) result(max_n)
    integer, intent(in) :: n_lines_b
    integer :: max_n
    max_n = 3 * (2 ** n_lines_b) - 2
  end function max_n_tris_c_tri

  pure elemental function max_n_tris_c_poly(n_lines_a, n_lines_b) result(max_n)
    integer, intent(in) :: n_lines_a, n_lines_b
    integer :: max_n
    max_n = n_lines_a * (2 ** n_lines_b) - 2
  end function max_n_tris_c_poly

  pure elemental function max_n_tets_c(n_planes_b) result(max_n)
    integer, intent(in) :: n_planes_b
    integer :: max_n
    max_n = 3 ** n_planes_b
  end function max_n_tets_c

  pure elemental function max_n_simplices_c_simplices(dim) result(max_n)
    integer, intent(in) :: dim
    integer :: max_n
    select case(dim)
      case(1)
        max_n = interval_buf_size
      case(2)
        max_n = tri_buf_size
      case(3)
        max_n = tet_buf_size
      case default
        max_n = 0
    end select
  end function max_n_simplices_c_simplices

  pure elemental function max_n_simplices_c_elements(dim, loc_a, loc_b) result(max_n)
    integer, intent(in) :: dim, loc_a, loc_b
    integer :: max_n
    max_n = 0
    if(loc_a == dim + 1 .and. loc_b == dim + 1) then
      max_n = max_n_simplices_c_simplices(dim)
    else if((loc_a == dim + 1 .and. loc_b == 2 ** dim) .or. (loc_b == dim + 1 .and. loc_a == 2 ** dim)) then
      if(dim == 2) max_n = max_n_tris_c_tri(4)
      if(dim == 3) max_n = max_n_tets_c(6)
    else if(loc_a == 2 ** dim .and. loc_b == 2 ** dim) then
      if(dim == 2) max_n = max_n_tris_c_poly(4, 4)
      if(dim == 3) max_n = 5 * max_n_tets_c(6)
    end if
  end function max_n_simplices_c_elements

  ! intersect_quads (Supermesh.F90:231-248): tris_c(2, 3, 62).
  subroutine intersect_quads(quad_a, quad_b, tris_c, n_tris_c)
    real(kind = real_kind), dimension(2, 4), intent(in) :: quad_a, quad_b
    real(kind = real_kind), dimension(:, :, :), intent(inout) :: tris_c
    integer, intent(out) :: n_tris_c
    real(kind = c_double), dimension(2, 4) :: a, b
    real(kind = c_double), dimension(2, 3, 62) :: c
    integer(kind = c_int) :: n
    a = quad_a
    b = quad_b
    call c_intersect_quads(a, b, c, n)
    n_tris_c = n
    if(n > 0) tris_c(:, :, :n) = c(:, :, :n)
  end subroutine intersect_quads

  ! intersect_hexes (Supermesh.F90:303-334): tets_c(3, 4, 5 * 729).
  subroutine intersect_hexes(hex_a, hex_b, tets_c, n_tets_c)
    real(kind = real_kind), dimension(3, 8), intent(in) :: hex_a, hex_b
    real(kind = real_kind), dimension(:, :, :), intent(inout) :: tets_c
    integer, intent(out) :: n_tets_c
    real(kind = c_double), dimension(3, 8) :: a, b
    real(kind = c_double), dimension(:, :, :), allocatable :: c
    integer(kind = c_int) :: n
    allocate(c(3, 4, 5 * 729))
    a = hex_a
    b = hex_b
    call c_intersect_hexes(a, b, c, n)
    n_tets_c = n
    if(n > 0) tets_c(:, :, :n) = c(:, :, :n)
  end subroutine intersect_hexes

  ! intersect_simplices (Supermesh.F90:141-160): intervals, triangles or
  ! tetrahedra by the dimension.
  subroutine intersect_simplices(simplex_a, simplex_b, simplices_c, n_simplices_c)
    real(kind = real_kind), dimension(:, :), intent(in) :: simplex_a, simplex_b
    real(kind = real_kind), dimension(:, :, :), intent(inout) :: simplices_c
    integer, intent(out) :: n_simplices_c
    real(kind = real_kind), dimension(2, 3, tri_buf_size) :: tc
    real(kind = real_kind), dimension(3, 4, tet_buf_size) :: ec
    real(kind = real_kind), dimension(2) :: ic
    n_simplices_c = 0
    select case(size(simplex_a, 1))
      case(1)
        call intersect_intervals(simplex_a(1, :), simplex_b(1, :), ic, n_simplices_c)
        if(n_simplices_c > 0) simplices_c(1, :, 1) = ic
      case(2)
        call intersect_tris_real(simplex_a, simplex_b, tc, n_simplices_c)
        if(n_simplices_c > 0) simplices_c(:, :, :n_simplices_c) = tc(:, :, :n_simplices_c)
      case(3)
        call intersect_tets_real(simplex_a, simplex_b, ec, n_simplices_c)
        if(n_simplices_c > 0) simplices_c(:, :, :n_simplices_c) = ec(:, :, :n_simplices_c)
    end select
  end subroutine intersect_simplices

  ! intersect_elements (Supermesh.F90:162-209): simplices, simplex x quad /
  ! hex either way round, quads or hexes, by the vertex counts.
  subroutine intersect_elements(element_a, element_b, elements_c, nelements_c)
    real(kind = real_kind), dimension(:, :), intent(in) :: element_a, element_b
    real(kind = real_kind), dimension(:, :, :), intent(inout) :: elements_c
    integer, intent(out) :: nelements_c
    integer :: dim, loc_a, loc_b
    dim = size(element_a, 1)
    loc_a = size(element_a, 2)
    loc_b = size(element_b, 2)
    nelements_c = 0
    if(loc_a == dim + 1 .and. loc_b == dim + 1) then
      call intersect_simplices(element_a, element_b, elements_c, nelements_c)
    else if(loc_a == dim + 1 .and. loc_b == 2 ** dim) then
      if(dim == 2) call intersect_tri_quad(element_a, element_b, elements_c, nelements_c)
      if(dim == 3) call intersect_tet_hex(element_a, element_b, elements_c, nelements_c)
    else if(loc_a == 2 ** dim .and. loc_b == dim + 1) then
      if(dim == 2) call intersect_tri_quad(element_b, element_a, elements_c, nelements_c)
      if(dim == 3) call intersect_tet_hex(element_b, element_a, elements_c, nelements_c)
    else if(loc_a == 2 ** dim .and. loc_b == 2 ** dim) then
      if(dim == 2) call intersect_quads(element_a, element_b, elements_c, nelements_c)
      if(dim == 3) call intersect_hexes(element_a, element_b, elements_c, nelements_c)
    end if
  end subroutine intersect_elements

  ! simplex_volume (Supermesh.F90:336-351).
  function simplex_volume(simplex) result(volume)
    real(kind = real_kind), dimension(:, :), intent(in) :: simplex
    real(kind = real_kind) :: volume
    select case(size(simplex, 1))
      case(1)
        volume = interval_size(simplex(1, :))
      case(2)
        volume = triangle_area(simplex)
      case(3)
        volume = tetrahedron_volume(simplex)
      case default
        volume = -huge(volume)
    end select
  end function simplex_volume

end module libsupermesh
