! Drives the Fortran interface the way the reference's serial tests do
! (src/tests/test_tri_intersector.F90:84-104, test_tet_intersector.F90:84-104):
! candidate pairs from the tree finder, then intersect_tris / intersect_tets
! per pair, summing triangle_area / tetrahedron_volume of the pieces.
! Usage: test_intersector_gpu mesh_a.bin mesh_b.bin
! Mesh file (stream, little endian): int32 nnodes, dim, nelements, loc;
! real(8) positions(dim, nnodes); int32 enlist(loc, nelements) (1-based).
! Prints: dim, number of candidate pairs, number of pieces, measure sum.
program test_intersector_gpu
  use libsupermesh
  implicit none
  character(len = 512) :: fa, fb
  real(kind = real_kind), dimension(:, :), allocatable :: pa, pb
  integer, dimension(:, :), allocatable :: ea, eb
  integer, dimension(:), allocatable :: indices, indptr
  real(kind = real_kind), dimension(2, 3, tri_buf_size) :: tris_c
  real(kind = real_kind), dimension(3, 4, tet_buf_size) :: tets_c
  real(kind = real_kind) :: total
  integer :: a, i, b, c, n, npieces, dim

  call get_command_argument(1, fa)
  call get_command_argument(2, fb)
  call read_mesh(trim(fa), pa, ea)
  call read_mesh(trim(fb), pb, eb)
  dim = size(pa, 1)
  allocate(indptr(size(ea, 2) + 1))
  call tree_intersection_finder(pa, ea, pb, eb, indices, indptr)
  if(len(libsupermesh_gpu_error()) > 0) then
    print "(a,a)", "error: ", libsupermesh_gpu_error()
    stop 1
  end if
  total = 0.0_real_kind
  npieces = 0
  do a = 1, size(ea, 2)
    do i = indptr(a), indptr(a + 1) - 1
      b = indices(i)
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
  print "(i0,1x,i0,1x,i0,1x,es25.17)", dim, size(indices), npieces, total

contains

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
