! Evaluates Fortran intrinsics used by the supermeshing hot path with the
! reference's compiler, for bit comparison with the CPU oracle.
! Input  (stream, little endian): int32 n, then n records of 6 real(8) (a, b).
! Output (stream): per record norm2(a), dot_product(a, b), spacing(a(1)),
!   the unit-normalised cross product a x b / norm2(a x b) (3 values) and
!   dot_product(b, unit cross) - dot_product(a, unit cross).
program probe_intrinsics
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  integer :: n, i, uin, uout
  real(kind = dp), dimension(3) :: a, b
  real(kind = dp), dimension(7) :: res
  character(len = 512) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit = uin, file = trim(fin), access = "stream", form = "unformatted", status = "old")
  open(newunit = uout, file = trim(fout), access = "stream", form = "unformatted", status = "replace")
  read(uin) n
  do i = 1, n
    read(uin) a, b
    call evaluate(a, b, res)
    write(uout) res
  end do
  close(uin)
  close(uout)
contains
  subroutine evaluate(a, b, res)
    real(kind = dp), dimension(3), intent(in) :: a, b
    real(kind = dp), dimension(7), intent(out) :: res
    real(kind = dp), dimension(3) :: c
    res(1) = norm2(a)
    res(2) = dot_product(a, b)
    res(3) = spacing(a(1))
    c = cross_unit(a, b)
    res(4:6) = c
    res(7) = dot_product(b, c) - dot_product(a, c)
  end subroutine evaluate
  pure function cross_unit(u, v) result(w)
    real(kind = dp), dimension(3), intent(in) :: u, v
    real(kind = dp), dimension(3) :: w
    w(1) = u(2) * v(3) - u(3) * v(2)
    w(2) = u(3) * v(1) - u(1) * v(3)
    w(3) = u(1) * v(2) - u(2) * v(1)
    w = w / norm2(w)
  end function cross_unit
end program probe_intrinsics
