! vl_magparser.f90 — command-line driver: runs an `exec run` control file on
! the gfx950 engine (the role of MAGPARSER.exe's control-file loop,
! src/controlparser.f90:341-593, for the operator path).
!
!   vl_magparser CONTROL_FILE [DEVICE]
!
! Prints the statement count, the ASSERT_ZERO failure count and the wall time;
! exits with status 3 when any ASSERT_ZERO failed.
program vl_magparser
  use iso_fortran_env, only: int64, output_unit, error_unit
  use vl_layer_ops, only: vl_fortran_init, vl_fortran_finalize
  use vl_control, only: run_control_file, assert_failures, statements_run
  implicit none
  character(len=4096) :: path, arg
  integer :: nargs, dev, ios
  integer(int64) :: t0, t1, rate
  nargs = command_argument_count()
  if (nargs < 1) then
     write(error_unit, '(A)') 'usage: vl_magparser CONTROL_FILE [DEVICE]'
     stop 2
  end if
  call get_command_argument(1, path)
  dev = 0
  if (nargs >= 2) then
     call get_command_argument(2, arg)
     read(arg, *, iostat=ios) dev
     if (ios /= 0) dev = 0
  end if
  call vl_fortran_init(dev)
  call system_clock(t0, rate)
  call run_control_file(trim(path))
  call system_clock(t1)
  write(output_unit, '(A,I0,A,I0,A,F0.3,A)') 'statements ', statements_run, ', ASSERT_ZERO failures ', &
       assert_failures, ', ', real(t1 - t0) / real(rate), ' s'
  call vl_fortran_finalize()
  if (assert_failures > 0) stop 3
end program vl_magparser
