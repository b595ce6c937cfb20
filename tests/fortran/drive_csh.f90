! Test driver for the Fortran drop-in boundary: the time loop of
! src/main2csh.f95:114-126 (derive2csh, ab2csh, history rotation), calling the
! reference's routine names, which resolve to tara_amd/fortran/csh_dropin.f90.
!
! usage: drive_csh IN OUT NSTEPS NU DT
!   IN / OUT: raw little-endian files: int32 Nx, int32 Ny, then omegak(Nh,Ny)
!   as complex(8), column-major.  The AB2 history starts at zero.
program drive_csh
  implicit none
  integer(kind=4) :: Nx, Ny, Nh, t, nsteps, u
  real(kind=8) :: nu, dt, time
  complex(kind=8), allocatable :: omegak(:,:), omegak_new(:,:), dt_omegak_old(:,:), dt_omegak_new(:,:)
  character(len=512) :: fin, fout, arg
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg); read(arg, *) nsteps
  call get_command_argument(4, arg); read(arg, *) nu
  call get_command_argument(5, arg); read(arg, *) dt
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) Nx, Ny
  Nh = Nx/2 + 1
  allocate(omegak(Nh,Ny), omegak_new(Nh,Ny), dt_omegak_old(Nh,Ny), dt_omegak_new(Nh,Ny))
  read(u) omegak
  close(u)
  dt_omegak_old = (0.0d0, 0.0d0)
  time = 0.0d0
  do t = 1, nsteps
    call derive2csh(Nx, Ny, Nh, nu, time, omegak, omegak_new, dt_omegak_old, dt_omegak_new)
    call ab2csh(Nx, Ny, Nh, nu, time, dt, omegak, omegak_new, dt_omegak_old, dt_omegak_new)
    dt_omegak_old = dt_omegak_new
    omegak = omegak_new
    time = time + dt
  end do
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) Nx, Ny
  write(u) omegak
  close(u)
end program drive_csh
