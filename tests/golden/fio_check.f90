! Fortran-side check of the msave / mload file formats
! (src/submodules/mlegs_scalar_io.f90), compiled and run by
! tests/test_field_io.py with the image's Fortran compiler.
!
!   fio_check DIR N1 N2 N3
!
! 1. reads DIR/c.bin (written by the library, global binary format) with
!    Fortran stream input: sizes, data, ln, offsets, space;
! 2. writes DIR/f.txt with the global ASCII output statements of msave
!    (format "(3(1X,I10))", then per row 2*N2 fields "1PE24.15E3" written
!    non-advancing, records closed and slices separated by list-directed
!    writes of '', then the list-directed trailer);
! 3. writes DIR/f.bin with msave's stream output statements;
! 4. reads DIR/c.txt (written by the library, global ASCII format) with the
!    input statements of mload (the size record, one formatted record of
!    2*N2 E fields per row, one skipped record between slices, the
!    list-directed trailer) and writes what it read to DIR/r.bin (stream).
program fio_check
  implicit none
  integer, parameter :: dp = selected_real_kind(15, 307)
  character(len=512) :: dir, arg
  integer :: n(3), m(3), u, i, j, k, off(3)
  complex(dp), allocatable :: a(:, :, :), b(:, :, :)
  real(dp), allocatable :: row(:)
  real(dp) :: ln
  character(len=3) :: sp
  character(len=64) :: efmt, rfmt

  call get_command_argument(1, dir)
  do i = 1, 3
    call get_command_argument(1 + i, arg)
    read(arg, *) n(i)
  enddo
  efmt = '(1PE24.15E3)'
  write(rfmt, '(a,i0,a)') '(', 2 * n(2), '(1PE24.15E3))'
  allocate(a(n(1), n(2), n(3)), b(n(1), n(2), n(3)), row(2 * n(2)))

  ! 1. stream input of the library's binary file
  open(newunit=u, file=trim(dir) // '/c.bin', form='unformatted', access='stream', status='old', action='read')
  read(u) m(1), m(2), m(3)
  if (any(m .ne. n)) error stop 'fio_check: sizes in c.bin differ'
  read(u) a
  read(u) ln, off(1), off(2), off(3)
  read(u) sp
  close(u)

  ! 2. global ASCII output
  open(newunit=u, file=trim(dir) // '/f.txt', status='replace')
  write(unit=u, fmt='(3(1X,I10))', advance='no') n(1), n(2), n(3)
  write(u, *) ''
  do k = 1, n(3)
    do i = 1, n(1)
      do j = 1, n(2)
        write(unit=u, fmt=trim(efmt), advance='no') real(a(i, j, k))
        write(unit=u, fmt=trim(efmt), advance='no') aimag(a(i, j, k))
      enddo
      if (i .lt. n(1)) write(u, *) ''
    enddo
    if (k .lt. n(3)) then
      write(u, *) ''
      write(u, *) ''
    endif
  enddo
  write(u, *) ''
  write(u, *) ln, off(1), off(2), off(3)
  write(u, *) sp
  close(u)

  ! 3. global binary output
  open(newunit=u, file=trim(dir) // '/f.bin', form='unformatted', access='stream', status='replace')
  write(u) n(1), n(2), n(3)
  write(u) a
  write(u) ln, off(1), off(2), off(3)
  write(u) sp
  close(u)

  ! 4. global ASCII input of the library's text file
  open(newunit=u, file=trim(dir) // '/c.txt', status='old')
  read(unit=u, fmt='(3(1X,I10))') m(1), m(2), m(3)
  if (any(m .ne. n)) error stop 'fio_check: sizes in c.txt differ'
  do k = 1, n(3)
    do i = 1, n(1)
      read(unit=u, fmt=trim(rfmt)) row
      do j = 1, n(2)
        b(i, j, k) = cmplx(row(2 * j - 1), row(2 * j), dp)
      enddo
    enddo
    if (k .lt. n(3)) read(u, *)
  enddo
  read(u, *) ln, off(1), off(2), off(3)
  read(u, *) sp
  close(u)
  open(newunit=u, file=trim(dir) // '/r.bin', form='unformatted', access='stream', status='replace')
  write(u) b
  write(u) ln, off(1), off(2), off(3)
  write(u) sp
  close(u)
end program fio_check
