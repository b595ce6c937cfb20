! File initial condition through the Fortran module: reads a one-field 2D
! checkpoint with tara_read_state2d, doubles it, writes it back with
! tara_write_state2d at step + 1 and time + 0.5.
! Usage: drive_ckpt in.ckpt out.ckpt Nx Ny
program drive_ckpt
  use iso_c_binding
  use tara_hip_mod
  implicit none
  character(len=512) :: inpath, outpath, arg
  integer :: nx, ny
  complex(c_double_complex), allocatable, target :: w(:, :)
  integer(c_long) :: step
  real(c_double) :: time
  call get_command_argument(1, inpath)
  call get_command_argument(2, outpath)
  call get_command_argument(3, arg)
  read(arg, *) nx
  call get_command_argument(4, arg)
  read(arg, *) ny
  allocate(w(nx / 2 + 1, ny))
  call tara_read_state2d(trim(inpath), nx / 2 + 1, ny, w, step, time)
  w = 2.0d0 * w
  call tara_write_state2d(trim(outpath), nx, ny, w, step + 1_c_long, time + 0.5d0)
end program
