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
    & c_f_pointer, c_associated

  implicit none

  private

  integer, parameter, public :: real_kind = c_double
  integer, parameter, public :: tri_buf_size = 22
  integer, parameter, public :: tet_buf_size = 81

  public :: tree_intersection_finder, rtree_intersection_finder, &
    & intersect_tris, intersect_tets, triangle_area, tetrahedron_volume, &
    & libsupermesh_gpu_error

  interface tree_intersection_finder
    module procedure tree_intersection_finder_csr_sparsity
  end interface tree_intersection_finder

  interface rtree_intersection_finder
    module procedure rtree_intersection_finder_csr_sparsity
  end interface rtree_intersection_finder

  interface intersect_tris
    module procedure intersect_tris_real
  end interface intersect_tris

  interface intersect_tets
    module procedure intersect_tets_real
  end interface intersect_tets

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
    function c_last_error() bind(c, name = "libsupermesh_gpu_last_error")
      import :: c_ptr
      type(c_ptr) :: c_last_error
    end function c_last_error
    function c_last_error_code() bind(c, name = "libsupermesh_gpu_last_error_code")
      import :: c_int
      integer(kind = c_int) :: c_last_error_code
    end function c_last_error_code
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
    call finder_csr(.false., positions_a, enlist_a, positions_b, enlist_b, map_ab_indices, &
      & map_ab_indptr)
  end subroutine tree_intersection_finder_csr_sparsity

  ! CSR candidate pairs, closed AABB overlap (Intersection_Finder.F90:621-643).
  subroutine rtree_intersection_finder_csr_sparsity(positions_a, enlist_a, positions_b, enlist_b, &
      & map_ab_indices, map_ab_indptr)
    real(kind = real_kind), dimension(:, :), intent(in) :: positions_a, positions_b
    integer, dimension(:, :), intent(in) :: enlist_a, enlist_b
    integer, dimension(:), allocatable, intent(out) :: map_ab_indices
    integer, dimension(:), intent(out) :: map_ab_indptr
    call finder_csr(.true., positions_a, enlist_a, positions_b, enlist_b, map_ab_indices, &
      & map_ab_indptr)
  end subroutine rtree_intersection_finder_csr_sparsity

  subroutine finder_csr(closed, positions_a, enlist_a, positions_b, enlist_b, map_ab_indices, &
      & map_ab_indptr)
    logical, intent(in) :: closed
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
    if(closed) then
      call c_rtree_set_input(nn_a, dim_a, ne_a, loc_a, nn_b, dim_b, ne_b, loc_b, pa, ea, pb, eb)
    else
      call c_tree_set_input(nn_a, dim_a, ne_a, loc_a, nn_b, dim_b, ne_b, loc_b, pa, ea, pb, eb)
    end if
    nind = 0
    if(c_last_error_code() == 0) then
      if(closed) then
        call c_rtree_query_output(nind)
      else
        call c_tree_query_output(nind)
      end if
    end if
    allocate(ind(max(nind, 1_c_long)), ptr(ne_a + 1))
    ptr = 0
    if(c_last_error_code() == 0) then
      if(closed) then
        call c_rtree_get_output(ne_a, nind, ind, ptr)
      else
        call c_tree_get_output(ne_a, nind, ind, ptr)
      end if
    end if
    allocate(map_ab_indices(nind))
    map_ab_indices = int(ind(:nind)) + 1
    map_ab_indptr = int(ptr) + 1
  end subroutine finder_csr

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

end module libsupermesh
