! Test driver for the device-resident mode from Fortran (module tara_hip_mod):
! a config-1 run kept on the GPU with the driver's tracer particles
! (src/main2csh.f95:84-126, 187-265), then a 3D run with the Energy.dat
! averages recorded on the device (src/main3com.f95:536-1071).
!
! usage: drive_resident IN OUT
!   IN : int32 np, then xp(np), yp(np) as real(8) (start positions)
!   OUT: xp, yp after 1001 steps; omega, ux, uy (64 x 64); 20 rows of the ten
!        3D averages (16^3, ABC flow, 1/450 diffusion, one row per step)
program drive_resident
  use iso_c_binding
  use tara_hip_mod
  implicit none
  real(c_double), parameter :: pi = 3.14159265358979323846d0
  integer(c_int), parameter :: n2 = 64, n3 = 16, nrec = 20
  integer(c_int) :: np
  integer :: u
  real(c_double), allocatable :: xp(:), yp(:), om(:,:), ux(:,:), uy(:,:), rows(:)
  type(c_ptr) :: csh, mhd
  type(tara_mhd_params) :: prm
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) np
  allocate(xp(np), yp(np))
  read(u) xp, yp
  close(u)

  ! 2D: shear layer (d = 3/128 written exactly), AB2, nu = 0.01, dt = 1e-3
  call tara_check(tara_csh_create(n2, n2, 0.01d0, 1.0d-3, 1_c_int, csh), 'tara_csh_create')
  call tara_check(tara_csh_ic_shear(csh, 2.0d0, 2.0d0, 25.0d0, 2.0d0, 0.0234375d0), 'tara_csh_ic_shear')
  call tara_check(tara_csh_advance_tracers(csh, 1001_c_int, 1.0d-3, 2.0d0 * pi, 2.0d0 * pi, 0_c_int, np, xp, yp), &
                  'tara_csh_advance_tracers')
  allocate(om(n2, n2), ux(n2, n2), uy(n2, n2))
  call tara_check(tara_csh_fields(csh, om, ux, uy), 'tara_csh_fields')
  call tara_check(tara_csh_destroy(csh), 'tara_csh_destroy')

  ! 3D: ABC flow, 1/450 diffusion, CS = 10, dt = 1e-3; the ten averages after every step
  prm%lx = 2.0d0 * pi
  prm%ly = 2.0d0 * pi
  prm%lz = 2.0d0 * pi
  prm%mu = 1.0d0 / 450.0d0
  prm%eta = 1.0d0 / 450.0d0
  prm%cs = 10.0d0
  prm%dt = 1.0d-3
  call tara_check(tara_mhd_create(n3, n3, n3, prm, mhd), 'tara_mhd_create')
  call tara_check(tara_mhd_ic_abc(mhd, 1.0d0, 0.1d0, 0.1d0, 1.0d0, 1.0d0, 1.0d0, 1.0d0, 1.0d0), 'tara_mhd_ic_abc')
  allocate(rows(10 * nrec))
  call tara_check(tara_mhd_record_energy(mhd, nrec, 1_c_int, rows), 'tara_mhd_record_energy')
  call tara_check(tara_mhd_destroy(mhd), 'tara_mhd_destroy')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) xp, yp, om, ux, uy, rows
  close(u)
end program drive_resident
