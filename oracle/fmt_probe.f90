! fmt_probe.f90 -- test infrastructure: the Fortran runtime's own output of
! the edit descriptors the reference's VTK files use (nse_dominik.f:208-235,
! set_levelset.f:131-198), for pinning the library's formatter on values
! the reference's runs do not produce (NaN, Inf, -0, three-digit exponents,
! subnormals, rounding ties).
!
! usage: fmt_probe values.bin n     (n doubles, native endian)
! prints each value with (e15.7), then the first three with (3e15.7), then
! 'DIMENSIONS ',3(1x,i3) and 'POINTS ',i9,' float' of the integers 64,
! 1000, 7 and 1234567890.
program fmt_probe
  implicit none
  character(len=512) :: fin, ns
  integer :: n, i
  double precision, allocatable :: v(:)
  call get_command_argument(1, fin)
  call get_command_argument(2, ns)
  read (ns, *) n
  allocate (v(n))
  open (10, file=trim(fin), access='stream', form='unformatted', status='old')
  read (10) v
  close (10)
  do i = 1, n
    write (6, "(e15.7)") v(i)
  end do
  write (6, "(3e15.7)") v(1), v(2), v(3)
  write (6, "('DIMENSIONS ',3(1x,i3))") 64, 1000, 7
  write (6, "('POINTS ',i9,' float')") 1234567890
end program
